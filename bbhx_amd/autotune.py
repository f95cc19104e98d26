"""Launch-shape autotuner for the fused direct likelihood of a ragged batch.

The persistent dense-band kernel (k_dir_like) has three free shape parameters:
the number of dense bins per work item (``chunk``), the threads per workgroup
(``block``) and the number of persistent workgroups (``grid``). The best
choice depends on the batch: many short bands favour small chunks (balance),
few long bands large ones (less per-item overhead). ``tune_direct`` times
every candidate on the caller's actual batch with device events on the
current stream and returns the fastest; results are cached per
(device, batch signature) for the life of the process and, if a path is given,
in a JSON file.
"""

from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

CHUNKS = (4096, 16384, 65536)
BLOCKS = (128, 256)
GRID_PER_CU = (4, 8, 16)

_cache: dict = {}


def _signature(lengths, num_modes, data_length, device_name):
    lengths = np.asarray(lengths)
    return (f"{device_name}|B={lengths.size}|M={num_modes}|N={data_length}|"
            f"Lsum={int(lengths.sum())}|Lmax={int(lengths.max())}")


def tune_direct(like, params, lengths, modes=None, reps=3, cache_path=None, candidates=None, progress=None):
    """Choose (chunk, block, grid) for ``like.get_ll_fused`` on this batch.

    like: a bbhx_amd.likelihood.Likelihood; params (12, B); lengths (B,);
    progress: optional callable(chunk, block, grid, ms) called after each candidate.
    Returns dict(chunk=, block=, grid=, ms=, table=[(chunk, block, grid, ms), ...]).
    """
    import torch

    from .likelihood import fused_param_rows

    dev = torch.cuda.current_device()
    props = torch.cuda.get_device_properties(dev)
    M = len(modes) if modes is not None else len(like.waveform_gen.amp_phase_gen.allowable_modes)
    sig = _signature(lengths, M, like.data_stream_length, props.name)
    if sig in _cache:
        return _cache[sig]
    if cache_path is not None and os.path.exists(cache_path):
        stored = json.loads(Path(cache_path).read_text())
        if sig in stored:
            _cache[sig] = stored[sig]
            return stored[sig]
    cu = props.multi_processor_count
    if candidates is None:
        candidates = [(c, b, g * cu) for c in CHUNKS for b in BLOCKS for g in GRID_PER_CU]
    rows, ts, te = fused_param_rows(like.waveform_gen, params, 0.0, 1.0, like.device)
    flags = torch.zeros(rows.shape[1], dtype=torch.int32, device=like.device)
    table = []
    ref = None
    plan = like.direct_plan(lengths, modes, min_chunk=min(c for c, _, _ in candidates))
    for chunk, block, grid in candidates:
        plan.set_launch(chunk, block, grid)
        flags.zero_()
        out = like.get_ll_fused_device(plan, rows, ts, te, flags, modes)  # warm-up
        torch.cuda.synchronize()
        if ref is None:
            ref = out.clone()
        elif not torch.allclose(out, ref, rtol=0.0, atol=1e-9 * (1.0 + float(ref.abs().max()))):
            raise RuntimeError(f"launch shape {(chunk, block, grid)} changed the result")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            like.get_ll_fused_device(plan, rows, ts, te, flags, modes)
        e1.record()
        torch.cuda.synchronize()
        table.append((chunk, block, grid, e0.elapsed_time(e1) / reps))
        if progress is not None:
            progress(*table[-1])
    del plan
    best = min(table, key=lambda t: t[3])
    res = dict(chunk=best[0], block=best[1], grid=best[2], ms=best[3], table=table)
    _cache[sig] = res
    if cache_path is not None:
        stored = json.loads(Path(cache_path).read_text()) if os.path.exists(cache_path) else {}
        stored[sig] = res
        Path(cache_path).write_text(json.dumps(stored, indent=1))
    return res
