"""Autotuner for the grouped contraction launches.

A plan's strided-GEMM launch is shaped by its tile variant (64x64x16,
64x64x16 with each wave 16 rows across the tile, 64x64x32, 64x128x16,
128x64x16, 128x128x16, and 64x64x16 / 64x64x32 with two K steps of loads in
flight), the number of tasks the
contraction ranges are split towards, and the longest contraction range of
one task (gecco_hip_plan_set_tuning).  The best choice depends on the block
shapes of the formula level and on the GPU, so the tuner builds the plan for
candidate settings and times its execute on the live buffers with HIP events
on the launch stream, one axis at a time (greedy coordinate search from the
defaults); a candidate replaces the current best only if it is faster by more
than `margin`.  No counterpart in the reference, whose batching constants are
fixed input keywords.
"""
AXES = (('variant', (1, 9, 6, 7, 0, 2, 3, 4)),
        ('target_tasks', (2048, 1024, 4096, 8192)),
        ('max_kchunk', (1024, 512, 2048)))
DEFAULT = {'variant': 1, 'target_tasks': 2048, 'max_kchunk': 1024}


def time_plan(plan, bufs, reps=5, warmup=2):
    """Milliseconds per execute (HIP events on the current stream)."""
    import torch
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        plan.execute(bufs, stream.cuda_stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        plan.execute(bufs, stream.cuda_stream)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_inplace(plan, bufs, axes=AXES, start=None, reps=5, margin=0.01, log=None):
    """The same search on one plan that keeps its host data
    (Plan.keep_host): every candidate is a Plan.retune, and the plan ends
    with the best settings.  Returns (options, ms, table); log(text), if
    given, is called after every candidate."""
    best_opts = dict(DEFAULT if start is None else start)
    if start is None and getattr(plan, '_elem', 8) == 4:
        best_opts['variant'] = 9          # the float32 plans' default tile
    plan.retune(**best_opts)
    best_ms = time_plan(plan, bufs, reps)
    table = [(dict(best_opts), best_ms)]
    for name, values in axes:
        for v in values:
            if best_opts.get(name) == v:
                continue
            opts = dict(best_opts)
            opts[name] = v
            plan.retune(**opts)
            ms = time_plan(plan, bufs, reps)
            table.append((opts, ms))
            if log:
                log(f'{opts}: {ms:.4f} ms (best {best_ms:.4f})')
            if ms < (1.0 - margin) * best_ms:
                best_opts, best_ms = opts, ms
    plan.retune(**best_opts)
    return best_opts, best_ms, table


def tune(build, bufs, axes=AXES, start=None, reps=5, margin=0.01):
    """build(options) -> finalized plan.  Returns (plan, options, ms, table)
    with table = [(options, ms), ...] of every candidate timed."""
    best_opts = dict(DEFAULT if start is None else start)
    best = build(best_opts)
    best_ms = time_plan(best, bufs, reps)
    table = [(dict(best_opts), best_ms)]
    for name, values in axes:
        for v in values:
            if best_opts.get(name) == v:
                continue
            opts = dict(best_opts)
            opts[name] = v
            plan = build(opts)
            ms = time_plan(plan, bufs, reps)
            table.append((opts, ms))
            if ms < (1.0 - margin) * best_ms:
                best, best_opts, best_ms = plan, opts, ms
            else:
                del plan
    return best, best_opts, best_ms, table
