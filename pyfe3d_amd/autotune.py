"""A small autotuner: time each candidate launch configuration of a GPU
step on the real inputs and keep the fastest, and remember the choice.

Used for the design sweep's batched banded solver (threads per workgroup
and the LDS panel width, threads of the triangular solves, the launch of
the narrow-band plates) and its persistent batched PCG (threads per
workgroup and the number of workgroups); the candidates and the timing
callable come from the caller, so the same search serves any kernel with
runtime launch parameters.
Timing: HIP events on the current stream around `repeats` calls, after one
untimed warm-up call per candidate.

Choices are kept in a table keyed by (kernel, element type, size bucket,
GPU architecture): the table shipped with the package
(``autotune_table.json``, measured on MI355X) and a user table
(``$PF3D_AUTOTUNE_CACHE``, default ``~/.cache/pyfe3d_amd/autotune.json``;
``off`` disables it) that receives new measurements.  A key found in
either table is used without timing anything.
"""
import json
import math
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
SHIPPED = os.path.join(HERE, 'autotune_table.json')


def tune(candidates, run, repeats=2, log=None):
    """candidates: list of dicts; run(candidate) launches the step.
    Returns (best candidate, [(candidate, ms per call), ...])."""
    if not candidates:
        raise ValueError('no candidate configuration')
    stream = torch.cuda.current_stream()
    table = []
    for c in candidates:
        run(c)  # warm-up (and first-use costs) outside the timing
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(repeats):
            run(c)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / repeats
        table.append((dict(c), ms))
        if log:
            log('autotune %s: %.3f ms' % (c, ms))
    best = min(table, key=lambda t: t[1])[0]
    return best, table


# ------------------------------------------------------------- the table
def arch(device=None):
    """The GPU's architecture name without feature suffixes, e.g. 'gfx950'."""
    name = torch.cuda.get_device_properties(device or 0).gcnArchName
    return name.split(':')[0]


def size_bucket(x):
    """Power-of-two bucket of a positive size: 2^round(log2 x)."""
    return 1 << max(0, int(round(math.log2(max(1, x)))))


def make_key(kernel, element, bucket, arch_name):
    """Table key; bucket is a tuple of ints (sizes already bucketed)."""
    return '%s|%s|%s|%s' % (kernel, element, 'x'.join(str(int(b)) for b in bucket), arch_name)


def user_path():
    p = os.environ.get('PF3D_AUTOTUNE_CACHE', '')
    if p.lower() == 'off':
        return None
    return p or os.path.join(os.path.expanduser('~'), '.cache', 'pyfe3d_amd', 'autotune.json')


def _read(path):
    try:
        with open(path) as fh:
            data = json.load(fh)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def lookup(key, shipped=SHIPPED):
    """The stored entry for key (user table first, then the shipped one),
    or None."""
    for path in (user_path(), shipped):
        if path:
            entry = _read(path).get(key)
            if entry is not None:
                return entry
    return None


def record(key, config, table=None):
    """Store config (and the timings that chose it) under key in the user
    table; best effort (an unwritable location only loses the memo)."""
    path = user_path()
    if not path:
        return False
    data = _read(path)
    data[key] = dict(config=config,
                     timings=[dict(c, ms=round(ms, 3)) for c, ms in (table or [])])
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + '.tmp%d' % os.getpid()
        with open(tmp, 'w') as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
        return True
    except OSError:
        return False
