"""Autotuner for the propagator's launch choices.

The library takes launch parameters per call (``tile_rows``: rows per
workgroup tile, 16 or 32; ``shots_per_block``: shots one workgroup steps
through, 0 = library default; ``march``: tiled, or z-marching recompute or
adjoint, in the gradient).  ``tune`` times short forward and gradient calls of the caller's
problem (same grid, shots and geometry, ``nt_trial`` steps) for each candidate
through the public API and keeps the fastest for each of the two calls.

Results are cached per problem shape, keyed by (padded grid, shots, receivers,
free surface, wavelet sharing, device name); the cache is process-local and,
when ``GBW_TUNE_CACHE`` names a file, also persisted there as JSON.  Access is
guarded by a lock.
"""

from __future__ import annotations

import json
import os
import sys
import threading
from typing import Dict, Iterable, List, Optional, Tuple

import torch

from .physics.wave import kernels as K

_lock = threading.Lock()
_cache: Dict[str, Tuple[int, ...]] = {}


def _key(nz, nx, nabc, n_shots, n_rcv, free_surface, shared) -> str:
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"
    return json.dumps([nz + 2 * nabc, nx + 2 * nabc, n_shots, n_rcv, bool(free_surface),
                       bool(shared), dev])


def _load_file():
    path = os.environ.get("GBW_TUNE_CACHE")
    if path and os.path.exists(path):
        try:
            with open(path) as fh:
                for k, v in json.load(fh).items():
                    _cache.setdefault(k, tuple(v))
        except (OSError, ValueError):
            pass


def _save_file():
    path = os.environ.get("GBW_TUNE_CACHE")
    if path:
        tmp = path + ".tmp"
        with open(tmp, "w") as fh:
            json.dump({k: list(v) for k, v in _cache.items()}, fh)
        os.replace(tmp, path)


def default_candidates(n_shots: int) -> List[Tuple[int, int]]:
    """(tile_rows, shots_per_block) pairs tried for each pass."""
    cands = [(16, 0), (32, 0)]
    for spb in (1, 2, 4, 8, 16):
        if spb < n_shots:
            cands.append((16, spb))
    return cands


def forward_candidates(n_shots: int) -> List[Tuple[int, int, int]]:
    """(tile_rows, shots_per_block, steps_per_launch) triples tried for the
    forward call (two-step launches always tile 16 rows)."""
    cands = [(tr, spb, 1) for tr, spb in default_candidates(n_shots)]
    for rows in (16, 24):
        cands += [(rows, spb, 2) for spb in (0, 4, 8, 16, 32, 64) if spb <= n_shots]
    return cands


def _timed(fn, repeats: int) -> float:
    """Best of ``repeats`` timed calls after one untimed call, in ms."""
    times = []
    for _ in range(repeats + 1):
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        pre = fn(None)
        start.record()
        fn(pre)
        end.record()
        torch.cuda.synchronize()
        times.append(start.elapsed_time(end))
    return min(times[1:])


def tune(args: dict, nt_trial: int = 60, candidates: Optional[Iterable[Tuple[int, int]]] = None,
         repeats: int = 2, verbose: bool = False) -> K.LaunchChoice:
    """Pick the fastest launch choice for the problem described by ``args``
    (the keyword arguments of ``acoustic_forward_kernel``; ``nt`` and ``src_v``
    are cut to ``nt_trial`` steps for the trials).  The forward call (with
    checkpoint stores) and the gradient call (recompute + adjoint, tiles or
    z-marching recompute) are timed and chosen separately.  Registers the
    result for this problem shape (``kernels.register_tuned``: later calls of
    the same shape use it unless a choice is pinned with
    ``kernels.set_launch_choice``) and returns it."""
    nz, nx, nabc = args["nz"], args["nx"], args["nabc"]
    src_v = args["src_v"]
    key = _key(nz, nx, nabc, args["src_n"], args["rcv_n"], args["free_surface"], src_v.dim() == 1)
    with _lock:
        if not _cache:
            _load_file()
        hit = _cache.get(key)
    skey = K.shape_key(nz, nx, nabc, args["src_n"], args["rcv_n"], args["free_surface"],
                       src_v.dim() == 1)
    if hit is not None and len(hit) == 6:
        choice = K.LaunchChoice(*hit)
        K.register_tuned(skey, choice)
        return choice
    nt = min(nt_trial, args["nt"])
    trial = dict(args, nt=nt, src_v=src_v[..., :nt].contiguous(),
                 checkpoint_segments=max(1, nt // 20))
    vp = args["vp"].detach().clone().requires_grad_(True)
    trial["vp"] = vp
    cands = list(candidates or default_candidates(args["src_n"]))

    def log(msg):
        if verbose:
            print("  autotune " + msg, file=sys.stderr, flush=True)

    # trials pin each candidate; whatever happens, the caller's pinned choice
    # (if any) is restored afterwards
    prev = K.get_launch_choice()
    try:
        best = _tune_trials(args, trial, vp, cands, log, repeats)
    finally:
        K.set_launch_choice(prev)
    with _lock:
        _cache[key] = best
        _save_file()
    choice = K.LaunchChoice(*best)
    K.register_tuned(skey, choice)
    return choice


def _tune_trials(args, trial, vp, cands, log, repeats):
    """Times the forward candidates, then the gradient candidates with the
    fastest forward; returns the LaunchChoice fields of the winner."""

    def fwd(pre):
        return K.acoustic_forward_kernel(**trial)["p"]

    best_f, best_ft = (16, 0, 2), float("inf")
    for tr, spb, spl in forward_candidates(args["src_n"]):
        K.set_launch_choice(K.LaunchChoice(tr, spb, steps_per_launch=spl))
        t = _timed(fwd, repeats)
        log(f"forward tile_rows={tr} shots_per_block={spb} steps_per_launch={spl}: {t:.2f} ms")
        if t < best_ft:
            best_f, best_ft = (tr, spb, spl), t

    def bwd(pre):
        if pre is None:
            vp.grad = None
            return K.acoustic_forward_kernel(**trial)["p"].sum()
        pre.backward()

    # the recompute and adjoint inside the gradient call follow the forward's
    # steps_per_launch and take their own tiles; the z-marching recompute
    # exists for one-step launches only, the z-marching adjoint (march 3) for
    # two-step launches: it steps at most 8 shots per workgroup, and the tiled
    # recompute beside it keeps the candidate's shots_per_block
    best_b, best_bt = (best_f[0], best_f[1], 1), float("inf")
    spl = best_f[2]
    bwd_cands = cands if spl == 1 else [(rows, spb) for rows in (16, 24)
                                        for spb in (0, 4, 8, 16, 32, 64)
                                        if spb <= args["src_n"]]
    for march in ((1, 2) if spl == 1 else (1, 3)):
        for tr, spb in bwd_cands:
            if march == 2 and tr != cands[0][0]:
                continue  # the marching recompute ignores tile_rows
            K.set_launch_choice(K.LaunchChoice(best_f[0], best_f[1], march, tr, spb, best_f[2]))
            t = _timed(bwd, repeats)
            log(f"gradient tile_rows={tr} shots_per_block={spb} march={march}: {t:.2f} ms")
            if t < best_bt:
                best_b, best_bt = (tr, spb, march), t
    return (best_f[0], best_f[1], best_b[2], best_b[0], best_b[1], best_f[2])


def batch_candidates() -> List[Tuple[int, int, int]]:
    """(shots_per_block, steps_per_launch, tile_rows) tried for a batched forward."""
    return [(0, 2, 0), (2, 2, 0), (4, 2, 0), (0, 1, 0), (0, 2, 24), (2, 2, 24), (4, 2, 24)]


def tune_batch(problems, nt_trial: int = 80, candidates=None, repeats: int = 2,
               verbose: bool = False):
    """Pick the batched-forward launch choice (shots per workgroup, steps per
    launch) for ``problems`` by timing each candidate on the same problems cut
    to ``nt_trial`` steps.  Returns a ``batch.BatchLaunch``; results are cached
    per batch shape like ``tune``."""
    from . import batch as B
    shapes = [(a["nz"] + 2 * a["nabc"], a["nx"] + 2 * a["nabc"], int(a["src_n"]), int(a["rcv_n"]),
               bool(a["free_surface"]), a["src_v"].dim() == 1) for a in problems]
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"
    key = json.dumps(["batch", shapes, dev])
    with _lock:
        if not _cache:
            _load_file()
        hit = _cache.get(key)
    if hit is not None and len(hit) == 3:
        return B.BatchLaunch(*hit)
    trial = []
    for a in problems:
        nt = min(nt_trial, int(a["nt"]))
        trial.append(dict(a, nt=nt, src_v=a["src_v"][..., :nt].contiguous()))
    best, best_t = None, float("inf")
    for spb, spl, rows in (candidates or batch_candidates()):
        bf = B.BatchedForward(trial, B.BatchLaunch(spb, spl, rows))
        times = []
        for _ in range(repeats + 1):
            start = torch.cuda.Event(enable_timing=True)
            end = torch.cuda.Event(enable_timing=True)
            start.record()
            bf.run()
            end.record()
            torch.cuda.synchronize()
            times.append(start.elapsed_time(end))
        t = min(times[1:])
        if verbose:
            print(f"  autotune batch shots_per_block={spb} steps_per_launch={spl} "
                  f"tile_rows={rows}: {t:.2f} ms", file=sys.stderr, flush=True)
        if t < best_t:
            best, best_t = (spb, spl, rows), t
        del bf
    with _lock:
        _cache[key] = best
        _save_file()
    return B.BatchLaunch(*best)


def batch_gradient_candidates() -> List[Tuple[int, int]]:
    """(shots_per_block, tile_rows) tried for a batched gradient."""
    return [(0, 0), (2, 0), (4, 0), (0, 24), (2, 24), (4, 24)]


def tune_batch_gradient(problems, nt_trial: int = 80, candidates=None, repeats: int = 2,
                        verbose: bool = False):
    """Pick the batched-gradient launch choice (shots per workgroup, tile shape)
    for ``problems`` by timing a misfit-gradient step of each candidate on the
    same problems cut to ``nt_trial`` steps.  Returns a ``batch.BatchLaunch``;
    results are cached per batch shape like ``tune``."""
    from . import batch as B
    shapes = [(a["nz"] + 2 * a["nabc"], a["nx"] + 2 * a["nabc"], int(a["src_n"]), int(a["rcv_n"]),
               bool(a["free_surface"]), a["src_v"].dim() == 1) for a in problems]
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"
    key = json.dumps(["batch_gradient", shapes, dev])
    with _lock:
        if not _cache:
            _load_file()
        hit = _cache.get(key)
    if hit is not None and len(hit) == 2:
        return B.BatchLaunch(shots_per_block=hit[0], tile_rows=hit[1])
    trial = []
    for a in problems:
        nt = min(nt_trial, int(a["nt"]))
        trial.append(dict(a, nt=nt, src_v=a["src_v"][..., :nt].contiguous()))
    best, best_t = None, float("inf")
    for spb, rows in (candidates or batch_gradient_candidates()):
        bg = B.BatchedGradient(trial, B.BatchLaunch(shots_per_block=spb, tile_rows=rows))
        d_obs = [torch.zeros_like(o["p"]) for o in bg.outputs]
        times = []
        for _ in range(repeats + 1):
            start = torch.cuda.Event(enable_timing=True)
            end = torch.cuda.Event(enable_timing=True)
            start.record()
            bg.misfit_step(d_obs)
            end.record()
            torch.cuda.synchronize()
            times.append(start.elapsed_time(end))
        t = min(times[1:])
        if verbose:
            print(f"  autotune batch gradient shots_per_block={spb} tile_rows={rows}: {t:.2f} ms",
                  file=sys.stderr, flush=True)
        if t < best_t:
            best, best_t = (spb, rows), t
        del bg, d_obs
    with _lock:
        _cache[key] = best
        _save_file()
    return B.BatchLaunch(shots_per_block=best[0], tile_rows=best[1])
