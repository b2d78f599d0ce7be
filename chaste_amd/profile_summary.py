"""Summarise a rocprofv3 kernel-trace CSV (run_kernel_trace.csv) into a short
markdown table: per-kernel call count and duration percentiles, plus one
step's timeline per precision.  Used to write the files under profiles/.

    python -m chaste_amd.profile_summary <kernel_trace.csv> <command> > profiles/<name>.md
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
        # bench.py runs warmup, a 5-step ranking pass, the timed steps, then an
        # event-instrumented pass: take a step from the middle third (timed).
        k = len(idx) // 3
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


if __name__ == '__main__':
    print(summarise(sys.argv[1], sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else 'Kernel trace'), end='')
