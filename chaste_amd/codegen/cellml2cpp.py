#!/usr/bin/env python3
"""CellML 1.0 -> C++ cardiac cell-model generator.

The reference builds its cell classes at configure time with the external
``chaste_codegen`` package (pinned ``>=0.10.0,<0.11.0`` in
``chaste_codegen.txt:1``; driven from ``cmake/Modules/ChasteMacros.cmake:65-99``)
from ``heart/src/odes/cellml/*.cellml``.  That package is not available here, so
this module restates the part of its job the monodomain path needs: it reads a
CellML 1.0 model, resolves the component connections, orders the algebraic
equations and prints one self-contained C++ header per model.

The header is precision-generic (``template <class R>``) and compiles both as
HIP device code (``__host__ __device__``) and as plain host C++ (the FP64 CPU
oracle includes the same header).  It provides

* ``evaluate<R, RL>(y, V, i_stim_area, p, cm, dy, g_inf, g_k)``: the
  right-hand side with V clamped (``dV/dt`` is not written; the tissue path
  clamps V exactly as ``AbstractCardiacCell::ComputeExceptVoltage`` does,
  ``heart/src/odes/AbstractCardiacCell.cpp:90-105``).  With ``RL = true`` the
  gating variables get the Rush-Larsen coefficients instead of a derivative:
  ``y_new = g_inf + (y - g_inf) * exp(g_k * dt)``
  (``heart/src/odes/AbstractRushLarsenCardiacCell.cpp:106-128``).
* ``dvdt<R>(...)``: ``dV/dt`` including the stimulus (single-cell runs, where
  V is integrated by the cell itself, ``AbstractCardiacCell::Compute``).
* ``iionic<R>(y, V, p, cm)``: total ionic current in uA/cm^2, the quantity
  ``AbstractCardiacCellInterface::GetIIonic`` returns
  (``heart/src/odes/AbstractCardiacCellInterface.hpp:227-256``).
* ``grl_step<R, ORDER>``: the generalized Rush-Larsen step of the
  reference's ``--grl1`` / ``--grl2`` cells (forward-difference diagonal
  Jacobian, ``heart/src/odes/AbstractGeneralizedRushLarsenCardiacCell.cpp``).
* ``be_linear``, ``be_residual``, ``be_jacobian``, ``be_step``: the
  backward-Euler form of the reference's ``*BackwardEuler`` cell classes
  (``heart/src/odes/AbstractBackwardEulerCardiacCell.hpp:116-128, 300-330``):
  the linear gates in closed form, the remaining states by Newton's method on
  the implicit step with an analytic Jacobian (symbolic differentiation of
  the model equations, as the reference's generator does with sympy).

Unit conventions follow the reference's tissue coupling: the stimulus passed
in is an *area* stimulus in uA/cm^2 (``GetIntracellularAreaStimulus``,
``heart/src/odes/AbstractCardiacCellInterface.cpp:106-119``).  Models whose
membrane currents are per capacitance (pA/pF, e.g. ten Tusscher 2006) divide it
by the tissue capacitance ``cm`` (uF/cm^2) and multiply the summed ionic current
by ``cm`` to return uA/cm^2.

A state variable is integrated by Rush-Larsen when its derivative is affine in
itself, ``dy/dt = a + b*y``, with ``a`` and ``b`` depending on no state other
than the membrane voltage (the classic Hodgkin-Huxley gate of Rush & Larsen,
IEEE TBME 25:389, 1978); every other state uses forward Euler.

Usage::

    python -m chaste_amd.codegen.cellml2cpp <model.cellml> <out.h> <namespace> \
        [--param comp.var] [--scale comp.var] ...

``--param`` turns a constant into a run-time parameter (default: its CellML
value); ``--scale`` multiplies a variable's defining value by a run-time factor
(default 1).  Parameters are laid out in the order given.
"""
import argparse
import os
import sys
import xml.etree.ElementTree as ET

CML = '{http://www.cellml.org/cellml/1.0#}'
MML = '{http://www.w3.org/1998/Math/MathML}'
CMETA = '{http://www.cellml.org/metadata/1.0#}'
CML11 = '{http://www.cellml.org/cellml/1.1#}'


def _tag(e):
    return e.tag.split('}')[1]


RDF = '{http://www.w3.org/1999/02/22-rdf-syntax-ns#}'
BQBIOL = '{http://biomodels.net/biology-qualifiers/}'
OXMETA = 'oxford-metadata#'

# CellML 1.0 unit prefixes and the built-in units this translator resolves
# (each built-in is its own dimension; only the ratios of like quantities --
# time, currents, capacitances -- are ever taken).
_PREFIX = {'yotta': 24, 'zetta': 21, 'exa': 18, 'peta': 15, 'tera': 12, 'giga': 9, 'mega': 6, 'kilo': 3,
           'hecto': 2, 'deka': 1, 'deca': 1, 'deci': -1, 'centi': -2, 'milli': -3, 'micro': -6, 'nano': -9,
           'pico': -12, 'femto': -15, 'atto': -18, 'zepto': -21, 'yocto': -24}
_BUILTIN = {'ampere', 'farad', 'second', 'volt', 'metre', 'meter', 'litre', 'liter', 'mole', 'kelvin', 'coulomb',
            'joule', 'siemens', 'ohm', 'gram', 'kilogram', 'newton', 'watt', 'hertz', 'tesla', 'weber', 'henry',
            'pascal', 'becquerel', 'gray', 'sievert', 'katal', 'lux', 'lumen', 'candela', 'radian', 'steradian',
            'celsius'}
_ALIAS = {'meter': 'metre', 'liter': 'litre'}


class CellMLModel:
    def __init__(self, path):
        root = ET.parse(path).getroot()
        self.name = root.get('name')
        self.vars = {}
        parent = {}

        def find(k):
            while parent[k] != k:
                parent[k] = parent[parent[k]]
                k = parent[k]
            return k

        # unit definitions (model level and inside components)
        self.unit_defs = {}
        for u in root.iter(CML + 'units'):
            if u.get('name') and u.get('base_units') != 'yes':
                self.unit_defs[u.get('name')] = [
                    (x.get('units'), x.get('prefix'), float(x.get('exponent', '1')),
                     float(x.get('multiplier', '1'))) for x in u.findall(CML + 'unit')]
            elif u.get('name'):
                self.unit_defs[u.get('name')] = 'base'
        self.components = []
        for comp in root.findall(CML + 'component'):
            cname = comp.get('name')
            self.components.append(cname)
            for v in comp.findall(CML + 'variable'):
                key = (cname, v.get('name'))
                self.vars[key] = dict(units=v.get('units'),
                                      init=v.get('initial_value'),
                                      cmeta=v.get(CMETA + 'id'))
                parent[key] = key
        for con in root.findall(CML + 'connection'):
            mc = con.find(CML + 'map_components')
            c1, c2 = mc.get('component_1'), mc.get('component_2')
            for mv in con.findall(CML + 'map_variables'):
                a = find((c1, mv.get('variable_1')))
                b = find((c2, mv.get('variable_2')))
                if a != b:
                    parent[a] = b

        self.defs = {}      # source var -> (component, rhs element)
        self.odes = {}      # state var -> (component, rhs element)
        self.time = None
        raw_defs = []
        for comp in root.findall(CML + 'component'):
            cname = comp.get('name')
            for math in comp.findall(MML + 'math'):
                for ap in math.findall(MML + 'apply'):
                    ch = list(ap)
                    assert _tag(ch[0]) == 'eq', 'only assignment equations supported'
                    lhs, rhs = ch[1], ch[2]
                    if _tag(lhs) == 'ci':
                        raw_defs.append(('def', (cname, lhs.text.strip()), cname, rhs))
                    else:
                        assert _tag(lhs) == 'apply' and _tag(list(lhs)[0]) == 'diff'
                        bvar = lhs.find(MML + 'bvar').find(MML + 'ci').text.strip()
                        ci = [c for c in lhs if _tag(c) == 'ci'][0].text.strip()
                        raw_defs.append(('ode', (cname, ci), cname, rhs))
                        self.time = (cname, bvar)
        # Every connected set has exactly one source: the member carrying an
        # initial value or an equation.  All members map onto that source.
        sources = {}
        for key, info in self.vars.items():
            if info['init'] is not None:
                sources.setdefault(find(key), key)
        for kind, key, cname, rhs in raw_defs:
            sources[find(key)] = key
        self.canon_map = {}
        for key in self.vars:
            r = find(key)
            self.canon_map[key] = sources.get(r, key)
        for kind, key, cname, rhs in raw_defs:
            src = self.canon_map[key]
            if kind == 'def':
                self.defs[src] = (cname, rhs)
            else:
                self.odes[src] = (cname, rhs)
        self.time = self.canon_map[self.time]
        # A derivative used inside an expression (dX/dt on some right-hand
        # side) becomes a variable of its own, defined by X's equation in X's
        # component, and the expression refers to it by an alias.
        for k, (c, rhs) in list(self.defs.items()) + list(self.odes.items()):
            self._replace_derivatives(rhs, c)
        # metadata: a cmeta:id names its variable's oxmeta term, unless an RDF
        # description of that id says which term (bqbiol:is oxford-metadata#...)
        self.meta = {}
        self.annotation = {}    # variable -> its oxmeta term, else its cmeta:id
        by_id = {}
        for key, info in self.vars.items():
            if info['cmeta']:
                self.meta[info['cmeta']] = self.canon_map[key]
                by_id[info['cmeta']] = self.canon_map[key]
                self.annotation.setdefault(self.canon_map[key], info['cmeta'])
        for d in root.iter(RDF + 'Description'):
            about = (d.get(RDF + 'about') or '').lstrip('#')
            if about not in by_id:
                continue
            for q in d.findall(BQBIOL + 'is'):
                res = q.get(RDF + 'resource') or ''
                if OXMETA in res:
                    self.meta.setdefault(res.split(OXMETA)[1], by_id[about])
                    self.annotation[by_id[about]] = res.split(OXMETA)[1]
        self.V = self.meta['membrane_voltage']
        self.stim = self.meta['membrane_stimulus_current']
        # rates per millisecond: a model in other time units has every
        # derivative divided by the length of its time unit in ms
        self.time_factor = self.unit_scale(self.vars[self.time]['units'], {'second': 1}) / 1e-3
        if abs(self.time_factor - 1.0) > 1e-12:
            for k, (c, rhs) in list(self.odes.items()):
                ap = ET.Element(MML + 'apply')
                ET.SubElement(ap, MML + 'divide')
                ap.append(rhs)
                cn = ET.SubElement(ap, MML + 'cn')
                cn.text = repr(self.time_factor)
                self.odes[k] = (c, ap)
        # state order: V first, then document order (component, variable)
        states = [k for k in self.vars if self.canon_map[k] == k and k in self.odes]
        states.remove(self.V)
        self.states = [self.V] + states

    def canon(self, cname, name):
        return self.canon_map[(cname, name)]

    def _replace_derivatives(self, el, cname):
        for i, ch in enumerate(list(el)):
            if _tag(ch) == 'apply' and _tag(list(ch)[0]) == 'diff':
                x = [c for c in ch if _tag(c) == 'ci'][0].text.strip()
                xk = self.canon_map[(cname, x)]
                dkey = (xk[0], 'd_%s_dt' % xk[1])
                if dkey not in self.defs:
                    self.vars[dkey] = dict(units=None, init=None, cmeta=None)
                    self.canon_map[dkey] = dkey
                    self.defs[dkey] = self.odes[xk]
                alias = 'd_%s__%s_dt' % xk
                if (cname, alias) not in self.canon_map:
                    self.vars[(cname, alias)] = dict(units=None, init=None, cmeta=None)
                    self.canon_map[(cname, alias)] = dkey
                ci = ET.Element(MML + 'ci')
                ci.text = alias
                el.remove(ch)
                el.insert(i, ci)
            else:
                self._replace_derivatives(ch, cname)

    def unit(self, name):
        """(factor, dimensions) of a units name: the quantity in SI-prefixed
        built-in units, dimensions a dict built-in -> exponent."""
        name = _ALIAS.get(name, name)
        if name == 'dimensionless':
            return 1.0, {}
        d = self.unit_defs.get(name)
        if d is None:
            if name in _BUILTIN:
                return 1.0, {name: 1}
            raise KeyError('unknown units ' + name)
        if d == 'base':
            return 1.0, {name: 1}
        f, dims = 1.0, {}
        for u, prefix, e, mult in d:
            fu, du = self.unit(u)
            p = 0 if prefix is None else (_PREFIX[prefix] if prefix in _PREFIX else int(prefix))
            f *= mult * (10.0 ** p * fu) ** e
            for b, x in du.items():
                dims[b] = dims.get(b, 0) + x * e
        return f, {b: x for b, x in dims.items() if abs(x) > 1e-12}

    def unit_scale(self, name, dims):
        """The factor of a units name whose dimensions must equal dims."""
        f, d = self.unit(name)
        if d != dims:
            raise NotImplementedError('units %s are not %s' % (name, dims))
        return f

    def capacitance(self):
        """The membrane capacitance variable (oxmeta membrane_capacitance,
        else a constant in farads in the membrane equation), or None."""
        k = self.meta.get('membrane_capacitance')
        if k is not None:
            return k
        for key, info in self.vars.items():
            if self.canon_map[key] == key and info['init'] is not None and info['units']:
                try:
                    if self.unit(info['units'])[1] == {'farad': 1}:
                        return key
                except KeyError:
                    pass
        return None

    def init(self, key):
        return float(self.vars[key]['init'])


def cident(key):
    return '%s__%s' % key


class Printer:
    """Translates MathML expression trees to C++ source text."""

    def __init__(self, model, names, is_const=None, v_affine=None):
        self.m = model
        self.names = names      # canonical var -> C expression
        self.is_const = is_const or (lambda key: False)
        # (k, c) when an expression is k V + c with constant k, c, else None
        self.v_affine = v_affine or (lambda el, cname: None)

    def const_expr(self, el, cname):
        """True when el depends on constants only (no state, parameter or time)."""
        t = _tag(el)
        if t == 'ci':
            return self.is_const(self.ref(cname, el))
        if t in ('cn', 'pi', 'exponentiale'):
            return True
        if t == 'apply':
            return all(self.const_expr(c, cname) for c in list(el)[1:]
                       if _tag(c) not in ('degree', 'logbase', 'bvar'))
        return False

    def const_scaled(self, el, cname):
        """(X source, k) when el is a product with constant factors k and
        other factors X, or X divided by a constant, else None."""
        if _tag(el) != 'apply':
            return None
        ch = list(el)
        op = _tag(ch[0])
        args = ch[1:]
        ev = Evaluator(self.m)
        if op == 'times' and len(args) >= 2:
            cs = [x for x in args if self.const_expr(x, cname)]
            xs = [x for x in args if not self.const_expr(x, cname)]
            if not cs or not xs:
                return None
            k = 1.0
            for x in cs:
                k *= float(ev.eval(x, cname, 0.0))
            src = ' * '.join(self.expr(x, cname) for x in xs)
            return '(' + src + ')', k
        if op == 'divide' and self.const_expr(args[1], cname) and not self.const_expr(args[0], cname):
            return self.expr(args[0], cname), 1.0 / float(ev.eval(args[1], cname, 0.0))
        return None

    def ref(self, cname, el):
        key = self.m.canon(cname, el.text.strip())
        return key

    def expr(self, el, cname):
        t = _tag(el)
        if t == 'ci':
            key = self.ref(cname, el)
            return self.names[key]
        if t == 'cn':
            return self.number(el)
        if t == 'pi':
            return 'R(3.14159265358979323846)'
        if t == 'exponentiale':
            return 'R(2.71828182845904523536)'
        if t == 'true':
            return 'true'
        if t == 'false':
            return 'false'
        if t == 'piecewise':
            pieces = []
            other = 'R(NAN)'
            for p in el:
                if _tag(p) == 'piece':
                    val, cond = list(p)
                    pieces.append((self.expr(cond, cname), self.expr(val, cname)))
                elif _tag(p) == 'otherwise':
                    other = self.expr(list(p)[0], cname)
            s = other
            for cond, val in reversed(pieces):
                s = '((%s) ? (%s) : (%s))' % (cond, val, s)
            return s
        if t == 'apply':
            ch = list(el)
            op = _tag(ch[0])
            args = [a for a in ch[1:] if _tag(a) not in ('degree', 'logbase', 'bvar')]
            a = [self.expr(x, cname) for x in args]
            if op == 'plus':
                return '(' + ' + '.join(a) + ')'
            if op == 'minus':
                if len(a) == 1:
                    return '(-%s)' % a[0]
                # exp(E) - 1 and 1 - exp(E) without cancellation for small E
                e0, e1 = _exp_arg(args[0]), _exp_arg(args[1])
                if e0 is not None and _is_one(args[1]):
                    return 'cmx_expm1(%s)' % self.expr(e0, cname)
                if e1 is not None and _is_one(args[0]):
                    return '(-cmx_expm1(%s))' % self.expr(e1, cname)
                return '(%s - %s)' % (a[0], a[1])
            if op == 'times':
                return '(' + ' * '.join(a) + ')'
            if op == 'divide':
                # constant denominators become a multiplication by their
                # reciprocal (folded at compile time); others go through
                # cmx_div (exact in double, fast reciprocal in float)
                if self.const_expr(args[1], cname):
                    return '(%s * (R(1) / %s))' % (a[0], a[1])
                return 'cmx_div(%s, %s)' % (a[0], a[1])
            if op == 'power':
                ex = args[1]
                if _tag(ex) == 'cn':
                    v = float(self.number_value(ex))
                    if v == int(v) and 1 <= v <= 4:
                        return 'cmx_ipow<%d>(%s)' % (int(v), a[0])
                return 'pow(%s, %s)' % (a[0], a[1])
            if op == 'root':
                deg = el.find(MML + 'degree')
                if deg is None:
                    return 'sqrt(%s)' % a[0]
                return 'pow(%s, R(1) / %s)' % (a[0], self.expr(list(deg)[0], cname))
            if op == 'exp':
                aff = self.v_affine(args[0], cname)
                if aff is not None:
                    # exponent affine in V: single precision folds it to one
                    # FMA in base 2 (cmx_exp_v), double keeps it as written
                    k, c = aff
                    return 'cmx_exp_v(%s, V, %r, %r)' % (a[0], k * LOG2E, c * LOG2E)
                scaled = self.const_scaled(args[0], cname)
                if scaled is not None:
                    # exponent k * X with constant k: log2(e) folds into k
                    x, k = scaled
                    return 'cmx_exp_c(%s, %s, %r)' % (a[0], x, k * LOG2E)
                return 'cmx_exp(%s)' % a[0]
            if op == 'ln':
                return 'cmx_log(%s)' % a[0]
            if op == 'log':
                return 'log10(%s)' % a[0]
            if op == 'abs':
                return 'fabs(%s)' % a[0]
            if op == 'floor':
                return 'floor(%s)' % a[0]
            if op == 'ceiling':
                return 'ceil(%s)' % a[0]
            rel = {'lt': '<', 'gt': '>', 'leq': '<=', 'geq': '>=', 'eq': '==', 'neq': '!='}
            if op in rel:
                return '(%s %s %s)' % (a[0], rel[op], a[1])
            if op == 'and':
                return '(' + ' && '.join(a) + ')'
            if op == 'or':
                return '(' + ' || '.join(a) + ')'
            if op == 'not':
                return '(!%s)' % a[0]
            raise NotImplementedError(op)
        raise NotImplementedError(t)

    @staticmethod
    def number_value(el):
        if el.get('type') == 'e-notation':
            parts = [x.strip() for x in el.itertext()]
            parts = [x for x in parts if x]
            return '%se%s' % (parts[0], parts[1])
        return el.text.strip()

    def number(self, el):
        v = self.number_value(el)
        f = float(v)
        return 'R(%s)' % repr(f)


LOG2E = 1.4426950408889634074


def _exp_arg(el):
    """E when el is exp(E), else None."""
    if _tag(el) == 'apply':
        ch = list(el)
        if _tag(ch[0]) == 'exp':
            return ch[1]
    return None


def _is_one(el):
    return _tag(el) == 'cn' and float(Printer.number_value(el)) == 1.0


class Evaluator:
    """Numeric evaluation of model expressions with V given, states at their
    initial values, parameters at their defaults and stimulus / time zero
    (used to locate removable singularities)."""

    def __init__(self, model):
        self.m = model

    def var(self, key, V, memo):
        m = self.m
        if key == m.V:
            return V
        if key == m.stim or key == m.time:
            return 0.0
        if key in memo:
            return memo[key]
        if key in m.odes or key not in m.defs:
            val = m.init(key)
        else:
            c, rhs = m.defs[key]
            val = self.eval(rhs, c, V, memo)
        memo[key] = val
        return val

    def eval(self, el, cname, V, memo=None):
        import math
        memo = {} if memo is None else memo
        t = _tag(el)
        if t == 'ci':
            return self.var(self.m.canon(cname, el.text.strip()), V, memo)
        if t == 'cn':
            return float(Printer.number_value(el))
        if t == 'pi':
            return math.pi
        if t == 'exponentiale':
            return math.e
        if t == 'piecewise':
            for p in el:
                if _tag(p) == 'piece':
                    val, cond = list(p)
                    if self.eval(cond, cname, V, memo):
                        return self.eval(val, cname, V, memo)
                elif _tag(p) == 'otherwise':
                    return self.eval(list(p)[0], cname, V, memo)
            return float('nan')
        if t == 'apply':
            ch = list(el)
            op = _tag(ch[0])
            a = [self.eval(x, cname, V, memo) for x in ch[1:] if _tag(x) not in ('degree', 'logbase', 'bvar')]
            if op == 'plus':
                return sum(a)
            if op == 'minus':
                return -a[0] if len(a) == 1 else a[0] - a[1]
            if op == 'times':
                r = 1.0
                for x in a:
                    r *= x
                return r
            if op == 'divide':
                return a[0] / a[1] if a[1] != 0 else (float('nan') if a[0] == 0 else math.copysign(float('inf'), a[0]))
            if op == 'power':
                return a[0] ** a[1]
            if op == 'exp':
                return math.exp(min(a[0], 700.0))
            if op == 'ln':
                return math.log(a[0]) if a[0] > 0 else float('nan')
            if op == 'abs':
                return abs(a[0])
            rel = {'lt': lambda x, y: x < y, 'gt': lambda x, y: x > y, 'leq': lambda x, y: x <= y,
                   'geq': lambda x, y: x >= y, 'eq': lambda x, y: x == y, 'neq': lambda x, y: x != y}
            if op in rel:
                return rel[op](a[0], a[1])
            if op == 'and':
                return all(a)
            if op == 'or':
                return any(a)
            if op == 'not':
                return not a[0]
        raise NotImplementedError(t)


# Removable singularities of GHK-type current and rate expressions, e.g.
# (V - c) / (exp(k (V - c)) - 1): at V = c numerator and denominator both
# vanish.  The reference's code generator (chaste_codegen, Chaste >= 2021.1)
# removes these automatically; here an expression that has one is evaluated
# by linear interpolation between V = c - dV and c + dV when |V - c| < dV,
# with dV chosen so that the exponent k (V - c) spans +-1e-7.
SINGULARITY_SPAN = 1e-7


def deps_of(el, cname, model, acc):
    """Canonical variables referenced by an expression."""
    if _tag(el) == 'ci':
        acc.add(model.canon(cname, el.text.strip()))
    for c in el:
        deps_of(c, cname, model, acc)
    return acc


class Generator:
    def __init__(self, model, params, scales):
        self.m = model
        self.params = params    # list of canonical keys (constants -> parameter)
        self.scales = scales    # list of canonical keys (defined value x factor)
        self.param_index = {}
        for i, k in enumerate(params + scales):
            self.param_index[k] = i
        self.fixed = []         # (variable, singular V, half width) of removed singularities
        self.deps = {}
        for k, (c, rhs) in model.defs.items():
            self.deps[k] = deps_of(rhs, c, model, set())
        self._state_deps = {}

    def is_const(self, key):
        """Constants and algebraic variables built from constants only."""
        m = self.m
        memo = self.__dict__.setdefault('_const_memo', {})
        if key in memo:
            return memo[key]
        if key in m.odes or key == m.stim or key == m.time or key in self.params or key in self.scales:
            r = False
        elif key in m.defs:
            memo[key] = False   # guards cycles
            r = all(self.is_const(d) for d in self.deps[key])
        else:
            r = m.vars[key]['init'] is not None
        memo[key] = r
        return r

    def state_deps(self, key, stack=()):
        """States a variable depends on, transitively (through algebraic defs)."""
        m = self.m
        if key in self._state_deps:
            return self._state_deps[key]
        if key in m.odes:
            r = {key}
        elif key == m.stim:
            r = set()
        elif key in m.defs:
            r = set()
            for d in self.deps[key]:
                r |= self.state_deps(d)
        else:
            r = set()
        self._state_deps[key] = r
        return r

    def singular_points(self, key):
        """(c, dV) for each removable singularity of definition key: a
        division whose numerator or denominator holds exp(E) - 1 (or
        1 - exp(E)) with E linear in V alone, and whose numerator and
        denominator both vanish at the root c of E."""
        m = self.m
        cname, rhs = m.defs[key]
        ev = Evaluator(m)
        found = []

        def v_only(el, c):
            ds = deps_of(el, c, m, set())
            return all(d == m.V or self.is_const(d) for d in ds)

        def exp_minus_one(el, c, out):
            if _tag(el) == 'apply':
                ch = list(el)
                if _tag(ch[0]) == 'minus' and len(ch) == 3:
                    for e, one in ((_exp_arg(ch[1]), ch[2]), (_exp_arg(ch[2]), ch[1])):
                        if e is not None and _is_one(one) and v_only(e, c):
                            out.append(e)
                for x in ch[1:]:
                    exp_minus_one(x, c, out)

        def walk(el, c):
            if _tag(el) != 'apply':
                for x in el:
                    walk(x, c)
                return
            ch = list(el)
            if _tag(ch[0]) == 'divide':
                num, den = ch[1], ch[2]
                cands = []
                exp_minus_one(num, c, cands)
                exp_minus_one(den, c, cands)
                for e in cands:
                    e0, e1, e2 = ev.eval(e, c, 0.0), ev.eval(e, c, 1.0), ev.eval(e, c, 2.0)
                    slope = e1 - e0
                    if slope == 0 or abs(e2 - e0 - 2 * slope) > 1e-9 * (abs(e0) + abs(slope)):
                        continue
                    vc = -e0 / slope
                    n0, d0 = ev.eval(num, c, vc), ev.eval(den, c, vc)
                    n1, d1 = ev.eval(num, c, vc + 1.0), ev.eval(den, c, vc + 1.0)
                    if abs(n0) <= 1e-9 * abs(n1) and abs(d0) <= 1e-9 * abs(d1):
                        pt = (vc, SINGULARITY_SPAN / abs(slope))
                        if all(abs(pt[0] - q[0]) > 1e-9 for q in found):
                            found.append(pt)
            for x in ch[1:]:
                walk(x, c)
        walk(rhs, cname)
        return found

    def v_affine(self, el, cname):
        """(k, c) when el depends on the voltage and constants only and
        equals k V + c (checked at five voltages to 1e-12), else None."""
        m = self.m
        ds = deps_of(el, cname, m, set())
        if m.V not in ds or not all(d == m.V or self.is_const(d) for d in ds):
            return None
        ev = Evaluator(m)
        pts = (-90.0, -47.5, 0.0, 12.25, 60.0)
        try:
            vals = [float(ev.eval(el, cname, v)) for v in pts]
        except (TypeError, ValueError, ZeroDivisionError, OverflowError):
            return None
        c = vals[2]
        k = (vals[4] - vals[0]) / (pts[4] - pts[0])
        for v, f in zip(pts, vals):
            if not abs(f - (k * v + c)) <= 1e-12 * (abs(k * v) + abs(c) + abs(f)) + 1e-300:
                return None
        return k, c

    def closure(self, roots):
        """Algebraic variables needed by roots, in evaluation order."""
        m = self.m
        order, seen = [], set()

        def visit(k):
            if k in seen:
                return
            seen.add(k)
            if k in m.defs and k != m.stim:
                for d in sorted(self.deps[k]):
                    visit(d)
                order.append(k)
        for r in roots:
            visit(r)
        return order

    # --- affine decomposition dy/dt = a + b*y for Rush-Larsen ---------------
    def affine(self, el, cname, y, printer):
        m = self.m
        t = _tag(el)
        if t == 'ci':
            k = m.canon(cname, el.text.strip())
            if k == y:
                return (None, 'R(1)')
            if y in self.state_deps(k) and k in m.defs:
                c2, rhs = m.defs[k]
                return self.affine(rhs, c2, y, printer)
            return (printer.expr(el, cname), None)
        if not self._depends(el, cname, y):
            return (printer.expr(el, cname), None)
        if t != 'apply':
            return None
        ch = list(el)
        op = _tag(ch[0])
        args = ch[1:]
        parts = [self.affine(x, cname, y, printer) for x in args]
        if any(p is None for p in parts):
            return None
        if op == 'plus':
            a, b = None, None
            for pa, pb in parts:
                a, b = _add(a, pa), _add(b, pb)
            return (a, b)
        if op == 'minus':
            if len(parts) == 1:
                return (_neg(parts[0][0]), _neg(parts[0][1]))
            return (_sub(parts[0][0], parts[1][0]), _sub(parts[0][1], parts[1][1]))
        if op == 'times':
            a, b = parts[0]
            for pa, pb in parts[1:]:
                if b is not None and pb is not None:
                    return None
                na = _mul(a, pa)
                nb = _add(_mul(a, pb), _mul(b, pa))
                a, b = na, nb
            return (a, b)
        if op == 'divide':
            (na, nb), (da, db) = parts
            if db is not None:
                return None
            return (_div(na, da), _div(nb, da))
        return None

    def rl_coeffs(self, rhs, cname, y, pr, aff):
        """(g_inf, g_k) source for a gate dy/dt = a + b y.  For the usual
        (y_inf - y) / tau form the steady state is y_inf itself and the rate
        -1/tau, without the a/b round trip."""
        if _tag(rhs) == 'apply' and _tag(list(rhs)[0]) == 'divide':
            num, den = list(rhs)[1:]
            if not self._depends(den, cname, y):
                part = self.affine(num, cname, y, pr)
                if part is not None and part[1] is not None:
                    na, nb = part
                    d = pr.expr(den, cname)
                    if nb == '(-R(1))':
                        return (na if na is not None else 'R(0)'), 'cmx_div(R(-1), %s)' % d
                    return ('cmx_div(-(%s), %s)' % (na or 'R(0)', nb)), 'cmx_div(%s, %s)' % (nb, d)
        a, b = aff
        return 'cmx_div(-(%s), %s)' % (a if a is not None else 'R(0)', b), b

    def _depends(self, el, cname, y):
        for k in deps_of(el, cname, self.m, set()):
            if k == y or y in self.state_deps(k):
                return True
        return False

    def generate(self, ns, source_name):
        m = self.m
        names = {}
        for k, info in m.vars.items():
            if m.canon_map[k] != k:
                continue
            names[k] = cident(k)
        for i, k in enumerate(m.states):
            names[k] = 'y[%d]' % i
        names[m.V] = 'V'
        names[m.stim] = 'i_stim'
        names[m.time] = 't'
        for k in self.params:
            names[k] = 'p[%d]' % self.param_index[k]
        pr = Printer(m, names, self.is_const, self.v_affine)

        constants = [k for k, info in m.vars.items()
                     if m.canon_map[k] == k and info['init'] is not None
                     and k not in m.odes and k not in self.params]
        state_ix = {k: i for i, k in enumerate(m.states)}

        # The stimulus arrives as an area density i_stim_area (uA/cm^2) with
        # the tissue capacitance cm (uF/cm^2), and the model's stimulus
        # variable takes it in its own units: a density per area as it is, a
        # current per capacitance as i_stim_area / cm, an absolute current as
        # i_stim_area / cm times the model's membrane capacitance (the current
        # per capacitance is what the conversion keeps).
        if abs(m.unit_scale(m.vars[m.V]['units'], {'volt': 1}) - 1e-3) > 1e-15:
            raise NotImplementedError('membrane voltage must be in millivolts')
        f_stim, d_stim = m.unit(m.vars[m.stim]['units'])

        def scaled(expr, k):
            return expr if abs(k - 1.0) < 1e-12 else '(%s) * R(%r)' % (expr, k)
        self.model_cap = None
        if d_stim == {'ampere': 1, 'metre': -2}:
            per_cap = False
            self.stim_expr = scaled('i_stim_area', 1e-2 / f_stim)            # 1 uA/cm^2 = 1e-2 A/m^2
        elif d_stim == {'ampere': 1, 'farad': -1}:
            per_cap = True
            self.stim_expr = scaled('i_stim_area / cm', 1.0 / f_stim)         # 1 uA/uF = 1 A/F
        elif d_stim == {'ampere': 1}:
            per_cap = True
            ck = m.capacitance()
            if ck is None:
                raise NotImplementedError('absolute stimulus current without a membrane capacitance')
            fc = m.unit_scale(m.vars[ck]['units'], {'farad': 1})
            self.model_cap = ck
            self.stim_expr = scaled('i_stim_area / cm', m.init(ck) * fc / f_stim)
        else:
            raise NotImplementedError('stimulus current in units %s' % m.vars[m.stim]['units'])

        def body_lines(roots, indent='    ', printer=None):
            printer = printer or pr
            lines = []
            for k in self.closure(roots):
                c, rhs = m.defs[k]
                e = printer.expr(rhs, c)
                for vc, dv in self.singular_points(k):
                    # the lambda's V shadows the clamped voltage
                    e = 'cmx_remove_singularity(V, R(%r), R(%r), [&](const R V) { return %s; })' % (vc, dv, e)
                    self.fixed.append((cident(k), vc, dv))
                if k in self.scales:
                    e = '(%s) * p[%d]' % (e, self.param_index[k])
                lines.append('%sconst R %s = %s;' % (indent, cident(k), e))
            return lines

        out = []
        w = out.append
        w('// Generated by chaste_amd/codegen/cellml2cpp.py from %s -- do not edit.' % source_name)
        w('// Model: %s' % m.name)
        w('#pragma once')
        w('#include "model_common.h"')
        w('')
        w('namespace cmx { namespace %s {' % ns)
        w('')
        w('constexpr int kNumStates = %d;' % len(m.states))
        w('constexpr int kVIndex = 0;')
        w('constexpr int kNumParams = %d;' % (len(self.params) + len(self.scales)))
        w('constexpr bool kPerCapacitanceUnits = %s;' % ('true' if per_cap else 'false'))
        w('')
        w('// State variables (CellML component__variable), V first then document order.')
        w('static const char* const kStateNames[kNumStates] = {')
        for k in m.states:
            w('    "%s",' % cident(k))
        w('};')
        w('static constexpr double kInitialState[kNumStates] = {')
        for k in m.states:
            w('    %r,  // %s' % (m.init(k), cident(k)))
        w('};')
        npar = len(self.params) + len(self.scales)
        w('static const char* const kParamNames[%d] = {' % max(npar, 1))
        for k in self.params:
            w('    "%s",' % cident(k))
        for k in self.scales:
            w('    "scale:%s",' % cident(k))
        if npar == 0:
            w('    "",')
        w('};')
        w('// The names the reference gives the same parameters (the CellML oxmeta term,')
        w('// else the variable\'s cmeta:id; empty when the file annotates neither).')
        w('static const char* const kParamAliases[%d] = {' % max(npar, 1))
        for k in self.params + self.scales:
            w('    "%s",' % m.annotation.get(k, ''))
        if npar == 0:
            w('    "",')
        w('};')
        w('static constexpr double kParamDefaults[%d] = {' % max(npar, 1))
        for k in self.params:
            w('    %r,' % m.init(k))
        for k in self.scales:
            w('    1.0,')
        if npar == 0:
            w('    0.0,')
        w('};')

        # Rush-Larsen classification
        rl = {}
        for k in m.states[1:]:
            c, rhs = m.odes[k]
            aff = self.affine(rhs, c, k, pr)
            if aff is None or aff[1] is None:
                continue
            # coefficient state dependencies: V only
            ok = True
            for el_str in aff:
                pass
            dep_states = set()
            for d in self._affine_deps(rhs, c, k):
                dep_states |= ({d} if d in m.odes else self.state_deps(d))
            dep_states.discard(k)
            if dep_states - {m.V}:
                ok = False
            if ok:
                rl[k] = aff
        w('// 1 = integrated by Rush-Larsen (affine gate, V-dependent coefficients only)')
        w('static constexpr int kIsGate[kNumStates] = {%s};' %
          ', '.join('1' if k in rl else '0' for k in m.states))
        w('')

        # evaluate
        w('// Right-hand side with V clamped.  Non-gates (and, with RL=false, every')
        w('// state) get dy[i]; with RL=true gate i gets g_inf[i], g_k[i] such that')
        w('// y_i(t+dt) = g_inf[i] + (y_i - g_inf[i]) * exp(g_k[i] * dt).')
        w('// i_stim_area is in uA/cm^2; cm is the tissue capacitance in uF/cm^2.')
        w('template <class R, bool RL>')
        w('CMX_HD inline void evaluate(const R* y, R V, R i_stim_area, const R* p, R cm,')
        w('                            R* dy, R* g_inf, R* g_k)')
        w('{')
        w('    CMX_MATH_USING')
        w('    (void)p; (void)cm; (void)g_inf; (void)g_k; (void)i_stim_area;')
        roots = set()
        for k in m.states[1:]:
            roots |= self.deps_rhs(k)
        for k in self.used_constants(roots, constants):
            w('    const R %s = R(%r);' % (cident(k), m.init(k)))
        w('    const R i_stim = %s;' % self.stim_expr)
        w('    (void)i_stim;')
        roots = set()
        for k in m.states[1:]:
            roots |= self.deps_rhs(k)
        for line in body_lines(sorted(roots)):
            w(line)
        for k in m.states[1:]:
            i = state_ix[k]
            c, rhs = m.odes[k]
            if k in rl:
                ginf, gk = self.rl_coeffs(rhs, c, k, pr, rl[k])
                w('    if (RL) {')
                w('        g_k[%d] = %s;' % (i, gk))
                w('        g_inf[%d] = %s;' % (i, ginf))
                w('    } else {')
                w('        dy[%d] = %s;' % (i, pr.expr(rhs, c)))
                w('    }')
            else:
                w('    dy[%d] = %s;' % (i, pr.expr(rhs, c)))
        w('}')
        w('')
        # dV/dt
        w('// dV/dt including the stimulus (single-cell integration).')
        w('template <class R>')
        w('CMX_HD inline R dvdt(const R* y, R V, R i_stim_area, const R* p, R cm)')
        w('{')
        w('    CMX_MATH_USING')
        w('    (void)p; (void)cm;')
        for k in self.used_constants(self.deps_rhs(m.V), constants):
            w('    const R %s = R(%r);' % (cident(k), m.init(k)))
        w('    const R i_stim = %s;' % self.stim_expr)
        w('    (void)i_stim;')
        for line in body_lines(sorted(self.deps_rhs(m.V))):
            w(line)
        c, rhs = m.odes[m.V]
        w('    return %s;' % pr.expr(rhs, c))
        w('}')
        w('')
        # iionic: -dV/dt with zero stimulus, in uA/cm^2
        cap_key = m.meta.get('Cm')
        w('// Total ionic current in uA/cm^2 (GetIIonic): minus dV/dt at zero stimulus,')
        w('// times the membrane capacitance that converts it to a current density.')
        w('template <class R>')
        w('CMX_HD inline R iionic(const R* y, R V, const R* p, R cm)')
        w('{')
        w('    CMX_MATH_USING')
        w('    (void)p; (void)cm;')
        used = self.used_constants(self.deps_rhs(m.V), constants)
        capname = None if per_cap else self._find_model_capacitance()
        if capname is not None and capname not in used:
            used.append(capname)
        for k in used:
            w('    const R %s = R(%r);' % (cident(k), m.init(k)))
        w('    const R i_stim = R(0);')
        w('    (void)i_stim;')
        for line in body_lines(sorted(self.deps_rhs(m.V))):
            w(line)
        c, rhs = m.odes[m.V]
        w('    const R dv = %s;' % pr.expr(rhs, c))
        if per_cap:
            w('    return -dv * cm;')
        else:
            # dV/dt = -(sum I)/C  ->  sum I = -C * dV/dt, C from the model
            w('    return -dv * %s;' % (cident(capname) if capname else 'R(1)'))
        w('}')
        w('')
        self._backward_euler(w, rl, names, constants, state_ix, per_cap, body_lines)
        self._generalized_rush_larsen(w)
        w('}}  // namespace cmx::%s' % ns)
        return '\n'.join(out) + '\n', rl

    # --- Backward Euler (the reference's BackwardEuler cell classes) --------
    def _backward_euler(self, w, rl, names, constants, state_ix, per_cap, body_lines):
        """be_linear / be_residual / be_jacobian.  States whose derivative is
        affine in themselves with V-only coefficients (the Rush-Larsen gates)
        are 'linear' and updated in closed form, y <- (y + a dt) / (1 - b dt);
        the others are solved together by Newton's method on
        g - y_old - dt f(g) = 0 with the analytic Jacobian
        (heart/src/odes/AbstractBackwardEulerCardiacCell.hpp:116-128,
        heart/src/odes/CardiacNewtonSolver.hpp:66-160)."""
        m = self.m
        nonlin = [k for k in m.states[1:] if k not in rl]
        lin = [k for k in m.states[1:] if k in rl]
        N = len(nonlin)
        nl_ix = {k: i for i, k in enumerate(nonlin)}
        gnames = dict(names)
        for k in nonlin:
            gnames[k] = 'g[%d]' % nl_ix[k]
        gpr = Printer(m, gnames, self.is_const, self.v_affine)
        w('// Backward Euler: the states solved by Newton (all but the linear gates).')
        w('constexpr int kNumNonlinear = %d;' % N)
        w('static constexpr int kNonlinear[%d] = {%s};' % (max(N, 1), ', '.join(str(state_ix[k]) for k in nonlin) or '0'))
        w('')
        w('// Linear states: y <- (y + a dt) / (1 - b dt) for dy/dt = a + b y, a and b')
        w('// functions of V (ComputeOneStepExceptVoltage, closed-form part).')
        w('template <class R>')
        w('CMX_HD inline void be_linear(R* y, R V, const R* p, R cm, R dt)')
        w('{')
        w('    CMX_MATH_USING')
        w('    (void)p; (void)cm; (void)V;')
        roots = set()
        for k in lin:
            roots |= self.deps_rhs(k)
        for k in self.used_constants(roots, constants):
            w('    const R %s = R(%r);' % (cident(k), m.init(k)))
        for line in body_lines(sorted(roots)):
            w(line)
        for k in lin:
            a, b = rl[k]
            w('    y[%d] = (y[%d] + (%s) * dt) / (R(1) - (%s) * dt);' % (state_ix[k], state_ix[k], a or 'R(0)', b))
        w('}')
        w('')
        roots = set()
        for k in nonlin:
            roots |= self.deps_rhs(k)
        consts = self.used_constants(roots, constants)
        w('// Newton residual of the nonlinear states at guess g (y holds the other')
        w('// states, linear ones already advanced): g - y_old - dt f(g).')
        w('template <class R>')
        w('CMX_HD inline void be_residual(const R* y, const R* g, R V, R i_stim_area, const R* p, R cm, R dt, R* res)')
        w('{')
        w('    CMX_MATH_USING')
        w('    (void)y; (void)p; (void)cm; (void)i_stim_area;')
        for k in consts:
            w('    const R %s = R(%r);' % (cident(k), m.init(k)))
        w('    const R i_stim = %s;' % self.stim_expr)
        w('    (void)i_stim;')
        for line in body_lines(sorted(roots), printer=gpr):
            w(line)
        for k in nonlin:
            c, rhs = m.odes[k]
            i = nl_ix[k]
            w('    res[%d] = g[%d] - y[%d] - dt * %s;' % (i, i, state_ix[k], gpr.expr(rhs, c)))
        w('}')
        w('')
        w('// Jacobian of be_residual in g: J[i*N + j] = delta_ij - dt d f_i / d g_j.')
        w('template <class R>')
        w('CMX_HD inline void be_jacobian(const R* y, const R* g, R V, R i_stim_area, const R* p, R cm, R dt, R* J)')
        w('{')
        w('    CMX_MATH_USING')
        w('    (void)y; (void)p; (void)cm; (void)i_stim_area;')
        for k in consts:
            w('    const R %s = R(%r);' % (cident(k), m.init(k)))
        w('    const R i_stim = %s;' % self.stim_expr)
        w('    (void)i_stim;')
        for line in body_lines(sorted(roots), printer=gpr):
            w(line)
        nlset = set(nonlin)
        dnames = {}
        for k in self.closure(sorted(roots)):
            c, rhs = m.defs[k]
            for j in sorted(self.state_deps(k) & nlset, key=lambda q: nl_ix[q]):
                d = self.derivative(rhs, c, j, gpr, dnames)
                if d is None:
                    continue
                for vc, dv in self.singular_points(k):
                    d = 'cmx_remove_singularity(V, R(%r), R(%r), [&](const R V) { return %s; })' % (vc, dv, d)
                if k in self.scales:
                    d = '(%s) * p[%d]' % (d, self.param_index[k])
                nm = 'd_%s__d%d' % (cident(k), nl_ix[j])
                dnames[(k, j)] = nm
                w('    const R %s = %s;' % (nm, d))
        for k in nonlin:
            c, rhs = m.odes[k]
            i = nl_ix[k]
            for j in nonlin:
                jj = nl_ix[j]
                d = self.derivative(rhs, c, j, gpr, dnames)
                diag = 'R(1)' if i == jj else 'R(0)'
                if d is None:
                    w('    J[%d] = %s;' % (i * N + jj, diag))
                else:
                    w('    J[%d] = %s - dt * %s;' % (i * N + jj, diag, d))
        w('}')
        w('')
        w('// One backward-Euler step of every state but V at the given V')
        w('// (ComputeOneStepExceptVoltage): linear states in closed form, then the')
        w('// nonlinear ones by Newton from their current values.  Returns the Newton')
        w('// iteration count, -1 if it did not converge.')
        w('template <class R>')
        w('CMX_HD inline int be_step(R* y, R V, R i_stim_area, const R* p, R cm, R dt)')
        w('{')
        w('    be_linear<R>(y, V, p, cm, dt);')
        if N == 0:
            # every state is linear in itself: the closed form is the step
            w('    (void)i_stim_area;')
            w('    return 0;')
            w('}')
            w('')
            return
        w('    R g[kNumNonlinear];')
        w('CMX_UNROLL')
        w('    for (int i = 0; i < kNumNonlinear; ++i) g[i] = y[kNonlinear[i]];')
        w('    const int its = cmx_newton<kNumNonlinear, R>(')
        w('        g, [&](const R* x, R* r) { be_residual<R>(y, x, V, i_stim_area, p, cm, dt, r); },')
        w('        [&](const R* x, R* J) { be_jacobian<R>(y, x, V, i_stim_area, p, cm, dt, J); });')
        w('CMX_UNROLL')
        w('    for (int i = 0; i < kNumNonlinear; ++i) y[kNonlinear[i]] = g[i];')
        w('    return its;')
        w('}')
        w('')

    # --- Generalized Rush-Larsen (first and second order) ------------------
    def _generalized_rush_larsen(self, w):
        """grl_step<R, ORDER>: every state (and V when it is free) advanced by
        y + f/a (exp(a dt) - 1), a the diagonal Jacobian entry by a forward
        difference of step 1e-8, forward Euler where |a| < 1e-8
        (heart/src/odes/AbstractGeneralizedRushLarsenCardiacCell.cpp:70-170,
        ode/src/solver/GRL1IvpOdeSolver.cpp:53-90 and GRL2IvpOdeSolver.cpp:
        53-118: the second-order form takes a half step, then re-evaluates
        each state's rate and diagonal term at the midpoint state with that
        state itself at its initial value)."""
        w('// Rates of the full system: dV/dt (0 when V is clamped) and the states.')
        w('template <class R>')
        w('CMX_HD inline void grl_rates(const R* y, R i_stim_area, const R* p, R cm, bool update_v, R* f)')
        w('{')
        w('    R gi[kNumStates], gk[kNumStates];')
        w('    evaluate<R, false>(y, y[0], i_stim_area, p, cm, f, gi, gk);')
        w('    f[0] = update_v ? dvdt<R>(y, y[0], i_stim_area, p, cm) : R(0);')
        w('}')
        w('')
        w('template <class R>')
        w('CMX_HD inline R grl_advance(R y0, R f, R a, R dt)')
        w('{')
        w('    using std::fabs; using std::exp;')
        w('    return fabs(a) < R(1e-8) ? y0 + f * dt : y0 + (f / a) * (exp(a * dt) - R(1));')
        w('}')
        w('')
        w('// One generalized Rush-Larsen step of length dt (ORDER 1 or 2); V is')
        w('// advanced too when update_v, else held.')
        w('template <class R, int ORDER>')
        w('CMX_HD inline void grl_step(R* y, R i_stim_area, const R* p, R cm, R dt, bool update_v)')
        w('{')
        w('    constexpr int S = kNumStates;')
        w('    const R delta = R(1e-8);')
        w('    R f[S], a[S], y0[S], t[S];')
        w('    const int first = update_v ? 0 : 1;')
        w('CMX_UNROLL')
        w('    for (int i = 0; i < S; ++i) y0[i] = y[i];')
        w('    grl_rates<R>(y, i_stim_area, p, cm, update_v, f);')
        w('CMX_UNROLL')
        w('    for (int i = 0; i < S; ++i) {')
        w('        if (i < first) continue;')
        w('        const R ys = y[i];')
        w('        y[i] = ys + delta;')
        w('        grl_rates<R>(y, i_stim_area, p, cm, update_v, t);')
        w('        a[i] = (t[i] - f[i]) / delta;')
        w('        y[i] = ys;')
        w('    }')
        w('    if constexpr (ORDER == 1) {')
        w('CMX_UNROLL')
        w('        for (int i = 0; i < S; ++i)')
        w('            if (i >= first) y[i] = grl_advance<R>(y0[i], f[i], a[i], dt);')
        w('        return;')
        w('    }')
        w('    R half = dt * R(0.5);')
        w('CMX_UNROLL')
        w('    for (int i = 0; i < S; ++i)')
        w('        if (i >= first) y[i] = grl_advance<R>(y0[i], f[i], a[i], half);')
        w('    R g[S];')
        w('CMX_UNROLL')
        w('    for (int i = 0; i < S; ++i) {')
        w('        if (i < first) continue;')
        w('        const R ym = y[i];')
        w('        y[i] = y0[i];')
        w('        grl_rates<R>(y, i_stim_area, p, cm, update_v, t);')
        w('        g[i] = t[i];')
        w('        y[i] = y0[i] + delta;')
        w('        grl_rates<R>(y, i_stim_area, p, cm, update_v, t);')
        w('        a[i] = (t[i] - g[i]) / delta;')
        w('        y[i] = ym;')
        w('    }')
        w('CMX_UNROLL')
        w('    for (int i = 0; i < S; ++i)')
        w('        if (i >= first) y[i] = grl_advance<R>(y0[i], g[i], a[i], dt);')
        w('}')
        w('')

    def derivative(self, el, cname, j, pr, dnames):
        """C++ source of d(el)/d(state j), or None when it is identically
        zero.  Algebraic variables that depend on j refer to their own
        derivative variables (dnames), emitted in evaluation order."""
        m = self.m
        t = _tag(el)
        if t == 'ci':
            k = m.canon(cname, el.text.strip())
            if k == j:
                return 'R(1)'
            if k in m.defs and k != m.stim and j in self.state_deps(k):
                return dnames.get((k, j))
            return None
        if t in ('cn', 'pi', 'exponentiale', 'true', 'false'):
            return None
        if t == 'piecewise':
            pieces, other, any_d = [], 'R(0)', False
            for pc in el:
                if _tag(pc) == 'piece':
                    val, cond = list(pc)
                    d = self.derivative(val, cname, j, pr, dnames)
                    any_d |= d is not None
                    pieces.append((pr.expr(cond, cname), d or 'R(0)'))
                elif _tag(pc) == 'otherwise':
                    d = self.derivative(list(pc)[0], cname, j, pr, dnames)
                    any_d |= d is not None
                    other = d or 'R(0)'
            if not any_d:
                return None
            out = other
            for cond, d in reversed(pieces):
                out = '((%s) ? (%s) : (%s))' % (cond, d, out)
            return out
        if t != 'apply':
            raise NotImplementedError(t)
        ch = list(el)
        op = _tag(ch[0])
        args = [a for a in ch[1:] if _tag(a) not in ('degree', 'logbase', 'bvar')]
        ds = [self.derivative(a, cname, j, pr, dnames) for a in args]
        if all(d is None for d in ds):
            return None
        ex = [pr.expr(a, cname) for a in args]
        if op == 'plus':
            return '(' + ' + '.join(d for d in ds if d is not None) + ')'
        if op == 'minus':
            if len(args) == 1:
                return '(-%s)' % ds[0]
            return _sub(ds[0], ds[1])
        if op == 'times':
            terms = []
            for i, d in enumerate(ds):
                if d is None:
                    continue
                terms.append('(' + ' * '.join([d] + [ex[l] for l in range(len(args)) if l != i]) + ')')
            return '(' + ' + '.join(terms) + ')'
        if op == 'divide':
            da, db = ds
            if db is None:
                return 'cmx_div(%s, %s)' % (da, ex[1])
            num = '(-(%s * %s))' % (ex[0], db) if da is None else '(%s * %s - %s * %s)' % (da, ex[1], ex[0], db)
            return 'cmx_div(%s, (%s * %s))' % (num, ex[1], ex[1])
        if op == 'power':
            da, db = ds
            if db is None:
                b = args[1]
                if _tag(b) == 'cn':
                    v = float(Printer.number_value(b))
                    if v == 1.0:
                        return da
                    if v == int(v) and 2 <= v <= 5:
                        return '(R(%r) * cmx_ipow<%d>(%s) * %s)' % (v, int(v) - 1, ex[0], da)
                return '(%s * pow(%s, %s - R(1)) * %s)' % (ex[1], ex[0], ex[1], da)
            base = 'pow(%s, %s)' % (ex[0], ex[1])
            inner = '%s * cmx_log(%s)' % (db, ex[0])
            if da is not None:
                inner += ' + cmx_div(%s * %s, %s)' % (ex[1], da, ex[0])
            return '(%s * (%s))' % (base, inner)
        if op == 'root':
            deg = el.find(MML + 'degree')
            if deg is None:
                return 'cmx_div(%s, (R(2) * sqrt(%s)))' % (ds[0], ex[0])
            dg = pr.expr(list(deg)[0], cname)
            return '((R(1) / %s) * pow(%s, R(1) / %s - R(1)) * %s)' % (dg, ex[0], dg, ds[0])
        if op == 'exp':
            return '(cmx_exp(%s) * %s)' % (ex[0], ds[0])
        if op == 'ln':
            return 'cmx_div(%s, %s)' % (ds[0], ex[0])
        if op == 'log':
            return 'cmx_div(%s, (%s * R(2.302585092994045684)))' % (ds[0], ex[0])
        if op == 'abs':
            return '((%s < R(0)) ? (-%s) : (%s))' % (ex[0], ds[0], ds[0])
        if op in ('floor', 'ceiling'):
            return None
        raise NotImplementedError(op)

    def _find_model_capacitance(self):
        m = self.m
        c, rhs = m.odes[m.V]
        for k in deps_of(rhs, c, m, set()):
            if m.vars[k]['units'] and 'F_per_cm2' in m.vars[k]['units']:
                return k
        return None

    def used_constants(self, roots, constants):
        need = set(roots)
        for k in self.closure(sorted(roots)):
            need |= self.deps[k]
        return [k for k in constants if k in need]

    def deps_rhs(self, k):
        c, rhs = self.m.odes[k]
        return deps_of(rhs, c, self.m, set())

    def _affine_deps(self, el, cname, y):
        """Non-y variables the affine coefficients reference (inlining defs of y)."""
        m = self.m
        acc = set()
        for k in deps_of(el, cname, m, set()):
            if k == y:
                continue
            if k in m.defs and y in self.state_deps(k):
                c2, rhs = m.defs[k]
                acc |= self._affine_deps(rhs, c2, y)
            else:
                acc.add(k)
        return acc


def _add(x, y):
    if x is None:
        return y
    if y is None:
        return x
    return '(%s + %s)' % (x, y)


def _neg(x):
    return None if x is None else '(-%s)' % x


def _sub(x, y):
    if y is None:
        return x
    if x is None:
        return '(-%s)' % y
    return '(%s - %s)' % (x, y)


def _mul(x, y):
    if x is None or y is None:
        return None
    return '(%s * %s)' % (x, y)


def _div(x, d):
    if x is None:
        return None
    return '(%s / %s)' % (x, d)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    ap.add_argument('cellml')
    ap.add_argument('out')
    ap.add_argument('namespace')
    ap.add_argument('--param', action='append', default=[])
    ap.add_argument('--scale', action='append', default=[])
    args = ap.parse_args(argv)
    model = CellMLModel(args.cellml)

    def key(s):
        c, v = s.split('.')
        return model.canon(c, v)
    gen = Generator(model, [key(s) for s in args.param], [key(s) for s in args.scale])
    src, rl = gen.generate(args.namespace, os.path.basename(args.cellml))
    with open(args.out, 'w') as f:
        f.write(src)
    print('%s: %d states, %d Rush-Larsen gates: %s' % (
        args.out, len(model.states), len(rl), ', '.join(cident(k) for k in rl)))
    for name, vc, dv in sorted(set(gen.fixed)):
        print('  removable singularity in %s at V = %.9g (+-%.3g mV)' % (name, vc, dv))


if __name__ == '__main__':
    main()
