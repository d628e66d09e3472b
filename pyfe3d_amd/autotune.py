"""A small autotuner: time each candidate launch configuration of a GPU
step on the real inputs and keep the fastest.

Used for the batched banded solver of the design sweep (threads per
workgroup and the LDS panel width, which decides how many workgroups fit
on a CU); the candidates and the timing callable come from the caller, so
the same search serves any kernel with runtime launch parameters.
Timing: HIP events on the current stream around `repeats` calls, after one
untimed warm-up call per candidate.
"""
import torch


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
