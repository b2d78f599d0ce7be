"""Summaries of rocprofv3 output for the files under profiles/.

Kernel trace (run_kernel_trace.csv): per-kernel call count and duration
percentiles, plus one step's timeline per precision:

    python -m chaste_amd.profile_summary <kernel_trace.csv> <command> > profiles/<name>.md

HBM traffic from two counter passes (FETCH_SIZE, WRITE_SIZE; KiB per
dispatch) over the bench and over the plain-copy calibration
(chaste_amd.hbm_calib), each counter scaled by bytes-copied / counter-value of
the largest copy:

    python -m chaste_amd.profile_summary traffic <fetch_dir> <write_dir> <json_out> > profiles/<name>.md
"""
import collections
import csv
import sys


def summarise(path, command, title):
    rows = list(csv.DictReader(open(path)))
    ev = sorted((int(r['Start_Timestamp']), int(r['End_Timestamp']),
                 r['Kernel_Name'].split('(')[0].replace('void ', '')) for r in rows)
    by = collections.defaultdict(list)
    for a, b, name in ev:
        by[name].append((b - a) / 1e3)
    out = ['# ' + title, '', 'Command (GPU box, repository root):', '', '    ' + command, '',
           'Kernel durations from the kernel trace (microseconds):', '',
           '| kernel | calls | avg | p10 | p50 | p90 | total ms |', '|---|---|---|---|---|---|---|']
    for name, v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        v = sorted(v)
        out.append('| `%s` | %d | %.2f | %.2f | %.2f | %.2f | %.2f |' % (
            name, len(v), sum(v) / len(v), v[len(v) // 10], v[len(v) // 2], v[9 * len(v) // 10], sum(v) / 1e3))
    for prec in ('double', 'float'):
        idx = [i for i, e in enumerate(ev) if 'cell_step' in e[2] and prec in e[2]]
        if len(idx) < 3:
            continue
        # a typical step: the one of median length among the middle third
        # (steps spanning a host synchronisation are the long tail)
        third = range(len(idx) // 3, 2 * len(idx) // 3)
        spans = sorted((ev[idx[j + 1]][0] - ev[idx[j]][0], j) for j in third if j + 1 < len(idx))
        k = spans[len(spans) // 2][1]
        i0, i1 = idx[k], idx[k + 1]
        out += ['', 'One %s step (start offset, duration, gap before; microseconds):' % ('FP64' if prec == 'double' else 'FP32-mixed'),
                '', '| t | duration | gap | kernel |', '|---|---|---|---|']
        t0 = ev[i0][0]
        prev = None
        for a, b, name in ev[i0:i1]:
            gap = (a - prev) / 1e3 if prev else 0.0
            out.append('| %.2f | %.2f | %.2f | `%s` |' % ((a - t0) / 1e3, (b - a) / 1e3, gap, name))
            prev = b
        out.append('| %.2f | | | next step |' % ((ev[i1][0] - t0) / 1e3))
    return '\n'.join(out) + '\n'


KIND_OF = {'cell_step_kernel': 'cell', 'refine_residual_kernel': 'refine', 'residual_kernel': 'residual',
           'cg_start_kernel': 'spmv', 'cg_iter_kernel': 'update', 'cg_last_kernel': 'last', 'finish_kernel': 'finish'}


def _kind(name):
    for key in ('refine_residual_kernel', 'cell_step_kernel', 'residual_kernel', 'cg_start_kernel',
                'cg_iter_kernel', 'cg_last_kernel', 'finish_kernel'):
        if key in name:
            return KIND_OF[key]
    return None


def _counters(path):
    rows = list(csv.DictReader(open(path)))
    by = collections.defaultdict(list)
    for r in rows:
        by[r['Kernel_Name'].split('(')[0].replace('void ', '')].append(float(r['Counter_Value']))
    return by


def traffic(fetch_dir, write_dir, json_out):
    """Per-kernel HBM bytes per dispatch (average over all dispatches, and over
    the dispatches that wrote anything, i.e. did their work)."""
    import json
    scale = {}
    for counter, d in (('FETCH_SIZE', fetch_dir), ('WRITE_SIZE', write_dir)):
        copies = _counters(d + '/calib_counter_collection.csv')['__amd_rocclr_copyBuffer']
        # the copies are 24 MiB and 1024 MiB, three each: calibrate on the largest
        scale[counter] = (1024 * 1024) / max(copies)   # bytes-per-KiB-of-counter ratio, in KiB per KiB
    f = _counters(fetch_dir + '/run_counter_collection.csv')
    w = _counters(write_dir + '/run_counter_collection.csv')
    out = {'scale': {'FETCH_SIZE': scale['FETCH_SIZE'], 'WRITE_SIZE': scale['WRITE_SIZE']}, 'mixed': {}, 'fp64': {}}
    lines = ['# HBM traffic per dispatch: TT06 slab, 442,401 nodes (bench workload)', '',
             'Counter passes (GPU box, repository root; one rocprofv3 run per counter):', '',
             '    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <fetch_dir> -o run -- python3 bench.py --steps 200 --warmup 20 --no-cpu-baseline --skip-propagating --no-kernel-events --no-graph --no-autotune --no-accuracy --skip-cfg1 --no-tables',
             '    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <fetch_dir> -o calib -- python3 -m chaste_amd.hbm_calib 24 1024',
             '    (the same two with WRITE_SIZE into <write_dir>)',
             '    python -m chaste_amd.profile_summary traffic <fetch_dir> <write_dir> profiles/hbm_traffic_slab442k.json', '',
             'Calibration (1024 MiB device copy): FETCH_SIZE x %.3f, WRITE_SIZE x %.3f to get bytes.' % (
                 scale['FETCH_SIZE'], scale['WRITE_SIZE']), '',
             '| kernel | dispatches | read MB (all) | write MB (all) | read MB (working) | write MB (working) |',
             '|---|---|---|---|---|---|']
    for name in sorted(f, key=lambda k: -sum(f[k])):
        kind = _kind(name)
        if kind is None or name not in w:
            continue
        fv, wv = f[name], w[name]
        n = min(len(fv), len(wv))
        rd = [x * 1024 * scale['FETCH_SIZE'] for x in fv[:n]]
        wr = [x * 1024 * scale['WRITE_SIZE'] for x in wv[:n]]
        work = [i for i in range(n) if wr[i] > 0.5 * max(wr)]
        avg_all = (sum(rd) + sum(wr)) / n
        avg_work = sum(rd[i] + wr[i] for i in work) / max(len(work), 1)
        prec = 'mixed' if ', float' in name or '<float>' in name else 'fp64'
        out[prec][kind] = {'bytes_per_dispatch': avg_all, 'bytes_per_working_dispatch': avg_work,
                           'dispatches': n, 'working_dispatches': len(work), 'kernel': name}
        lines.append('| `%s` | %d | %.2f | %.2f | %.2f | %.2f |' % (
            name, n, sum(rd) / n / 1e6, sum(wr) / n / 1e6,
            sum(rd[i] for i in work) / max(len(work), 1) / 1e6, sum(wr[i] for i in work) / max(len(work), 1) / 1e6))
    json.dump(out, open(json_out, 'w'), indent=1)
    return '\n'.join(lines) + '\n'


if __name__ == '__main__':
    if sys.argv[1] == 'traffic':
        print(traffic(sys.argv[2], sys.argv[3], sys.argv[4]), end='')
    else:
        print(summarise(sys.argv[1], sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else 'Kernel trace'), end='')
