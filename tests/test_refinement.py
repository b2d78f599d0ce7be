"""Mixed precision's FP64 refinement: an FP32 CG solve whose result is
accepted only when its FP64 true residual passes the reference's test
(||D^-1 r|| <= rtol ||D^-1 b||, linalg/src/LinearSystem.cpp:833-837), else
corrected by a second solve on that residual.  The true residual comes either
from the first solve's last budgeted CG iteration (folded, the default) or
from a kernel of its own (separate); both decide the same test.

* A 2-D TT06 Rush-Larsen sheet with a wave crossing it, graph replay on,
  printed every 50 steps: folded and separate each within 0.5 mV of the
  oracle's FP64 run at every printed step and 0.1 ms in activation time, and
  within 0.01 mV of each other (two solves that pass the same test).
* At rtol 1e-10 an FP32 recursion cannot certify upstroke steps: the
  true residual fails inside folded steps and the host resumes the
  refinement solve from the residual the folded iteration left; folded and
  separate then agree within 1e-3 mV.
* The separate path never folds; the folded path folds every step of a batch
  budgeted at least one iteration.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = (2, 60, 20, 0)


def sheet(cmx, separate, rtol=1e-8, **kw):
    coords, elems = cmx.slab_mesh(*DIMS, 0.01)
    stim_index = np.where(coords[:, 0] < 0.02, 0, -1).astype(np.int32)
    return cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), cmx.TT06_EPI, cmx.RUSH_LARSEN,
                          stimuli=[(-400000.0, 0.0, 0.5)], stim_index=stim_index, precision=cmx.MIXED,
                          ksp_rtol=rtol, ksp_atol=0.0, refine_sweeps=1, extrapolate=2, use_graph=True,
                          separate_refinement=separate, **kw)


def run(m, steps, every):
    m.set_activation(0.0, 1, 0.0)
    out = []
    for j in range(steps // every):
        m.step(j * every * 0.01, every)
        out.append(m.voltage())
    return np.array(out)


def test_folded_and_separate_refinement_match_oracle(oracle):
    from chaste_amd import cmx
    print('case refinement: folded vs separate vs oracle', flush=True)
    steps, every = 1200, 50
    got, act, st = {}, {}, {}
    for sep in (False, True):
        m = sheet(cmx, sep)
        got[sep] = run(m, steps, every)
        act[sep] = m.activation()[4]
        st[sep] = m.stats()
        m.close()
    n = got[False].shape[1]
    ref = oracle.mono_run(*DIMS, model=oracle.TT06_EPI, rush_larsen=True, sigma=(1.75, 1.75, 1.75),
                          stim_mag=-400000.0, stim_dur=0.5, stim_xlim=0.02, t_end=steps * 0.01, ksp_rtol=1e-12,
                          trace_nodes=np.arange(n), act_threshold=0.0)
    rp = ref['trace'][every - 1::every]
    for sep in (False, True):
        both = (act[sep] >= 0) & (ref['act'] >= 0)
        one = (act[sep] >= 0) != (ref['act'] >= 0)
        dv = np.abs(got[sep] - rp).max()
        print('separate=%d: max |dV| %.3e mV, act err %.3e ms, activated %d, stats %s' % (
            sep, dv, np.abs(act[sep][both] - ref['act'][both]).max(), both.sum(),
            {k: st[sep][k] for k in ('steps', 'cg_iterations', 'refine_sweeps', 'resumes', 'folded_refinements')}))
        assert both.sum() > n // 2
        assert dv <= 0.5
        assert np.all(np.maximum(act[sep], ref['act'])[one] >= steps * 0.01 - 0.1)
        assert np.abs(act[sep][both] - ref['act'][both]).max() <= 0.1
    d = np.abs(got[False] - got[True]).max()
    print('folded vs separate max |dV| %.3e mV' % d)
    assert d <= 0.01
    assert st[True]['folded_refinements'] == 0
    assert st[False]['folded_refinements'] > steps // 2
    assert st[False]['refine_sweeps'] == st[True]['refine_sweeps'] == steps


def test_folded_refinement_resumes_refinement_solve():
    """rtol 1e-10 with a second stimulus at the far edge at 8 ms: quiet steps
    pass the true-residual test inside folded batches, upstroke steps fail it
    and are resumed.  A 50-step call whose folded count stops short of 50
    with a resume is a folded batch cut by a resumed step."""
    from chaste_amd import cmx
    print('case refinement: resumed after a folded step', flush=True)
    coords, elems = cmx.slab_mesh(*DIMS, 0.01)
    stim_index = np.where(coords[:, 0] < 0.02, 0, np.where(coords[:, 0] > 0.58, 1, -1)).astype(np.int32)
    got, chunks = {}, {}
    for sep in (False, True):
        m = cmx.Monodomain(coords, elems, np.array([1.75, 1.75]), cmx.TT06_EPI, cmx.RUSH_LARSEN,
                           stimuli=[(-400000.0, 0.0, 0.5), (-400000.0, 8.0, 0.5)], stim_index=stim_index,
                           precision=cmx.MIXED, ksp_rtol=1e-10, ksp_atol=0.0, refine_sweeps=1, extrapolate=2,
                           use_graph=True, separate_refinement=sep)
        prev, out, ch = m.stats(), [], []
        for j in range(24):
            m.step(j * 0.5, 50)
            out.append(m.voltage())
            s = m.stats()
            ch.append((s['folded_refinements'] - prev['folded_refinements'], s['resumes'] - prev['resumes']))
            prev = s
        m.close()
        got[sep], chunks[sep] = np.array(out), ch
        print('separate=%d (folded, resumes) per 50 steps: %s' % (sep, ch), flush=True)
    d = np.abs(got[False] - got[True]).max()
    print('folded vs separate at rtol 1e-10: max |dV| %.3e mV' % d)
    assert any(0 < f < 50 and r > 0 for f, r in chunks[False])
    assert all(f == 0 for f, _ in chunks[True])
    assert d <= 1e-3
