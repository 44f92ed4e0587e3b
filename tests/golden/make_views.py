"""Regenerate tests/golden/views.json: the deep views fixed by the seeded zoom walk.

Two walk rules.  "boundary" (oracle/oracle.c, orc_zoom_walk) starts from the reference's full view
(centre -0.5+0i, width 4, mandelbrot.cpp:13-14), renders res x res at the
current width, and moves the centre to a uniformly random escaped sample with
count >= 0.9 N (splitmix64(seed)), preferring samples next to a non-escaped sample
(the set's boundary), then samples on a band edge; the width is divided by 10
until the next width would fall below the target, which is then used as is.  A
depth whose render has no escaped sample backs the walk off one step, where another
candidate is drawn from the same splitmix64 stream.  Past a depth where the
iteration budget no longer resolves the set's boundary this rule settles on a
smooth level curve of the escape count, and below it every sample of a frame has
one count (at 1e-100 with N 100000 every sample escapes at 99999).

"detail" (detail_walk below, over the oracle's centre render) keeps to structure:
among the escaped samples with count >= 0.9 N (or >= 0.9 of the greatest escaped
count when none reaches that) it draws uniformly, by splitmix64(seed), from those
whose 5 x 5 neighbourhood holds the most distinct counts; a render with one count
only, or with no escaped sample, backs the walk off one step to draw again.  The
centre moves by the drawn sample's offset in mpmath arithmetic of 3.33 bits per
decimal digit of the target width plus 200 bits.

Usage: python tests/golden/make_views.py [name ...]   (default: all)
"""
import json
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402

# name: (seed, N, target width, walk resolution, rule)
SPECS = {
    "t11": (11, 3000, "1e-16", 128, "boundary"),
    "t13": (13, 8000, "1e-25", 128, "boundary"),
    "c2_seed1": (1, 20000, "1e-30", 256, "boundary"),
    "c3_seed1": (1, 50000, "1e-200", 256, "boundary"),
    "c4_seed1": (1, 60000, "1e-40", 256, "boundary"),
    "c5_seed1": (1, 100000, "1e-100", 256, "boundary"),
    "c3_detail1": (1, 50000, "1e-200", 256, "detail"),
    "c5_detail1": (1, 100000, "1e-100", 256, "detail"),
}
# the learned skip's training centres: the detail walk of seeds 2 ... 13 (at
# 64 x 64 renders, 16 times cheaper than the configurations' 256 x 256)
for _s in range(2, 14):
    SPECS["c5_train%d" % _s] = (_s, 100000, "1e-100", 64, "detail")

_M64 = 2**64 - 1


def _splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


def _fmt(x):
    return mp.nstr(x, int(mp.mp.dps), min_fixed=-mp.inf, max_fixed=mp.inf)


def detail_walk(seed, N, target_width, res, win=2, nthreads=0, log=None):
    """The "detail" rule (module docstring); returns (re, im, steps) as strings."""
    texp = -int(mp.floor(mp.log10(mp.mpf(target_width))))
    mp.mp.prec = int(texp * 3.33) + 200
    tgt = mp.mpf(target_width)
    cre, cim, w = mp.mpf(-0.5), mp.mpf(0), mp.mpf(4)
    state, k, back, retries = seed, 0, [], 0
    while w / 10 >= tgt:
        counts = O.center_render(_fmt(cre), _fmt(cim), mp.nstr(w, 20), res, N, nthreads)
        esc = counts < N
        flat = len(np.unique(counts)) == 1
        if flat or not esc.any():
            if not back or retries >= 64:
                break
            cre, cim, w, k = back.pop()
            retries += 1
            continue
        thr = int(np.ceil(0.9 * N))
        cand = esc & (counts >= thr)
        if not cand.any():
            cand = esc & (counts >= int(np.ceil(0.9 * counts[esc].max())))
        P = np.pad(counts, win, mode="edge")
        S = np.sort(np.stack([P[win + dr: win + dr + res, win + dc: win + dc + res]
                              for dr in range(-win, win + 1) for dc in range(-win, win + 1)]), axis=0)
        score = np.where(cand, 1 + (np.diff(S, axis=0) != 0).sum(axis=0), 0)
        idx = np.flatnonzero(score.ravel() == score.max())
        state, rnd = _splitmix64(state)
        r, c = divmod(int(idx[(rnd * len(idx)) >> 64]), res)
        back.append((cre, cim, w, k))
        sz = w / res
        cre, cim, w, k = cre + sz * (c - res // 2), cim + sz * (res // 2 - r - 1), w / 10, k + 1
        if log:
            log("  step %d: %d distinct counts, %d..%d" % (k, len(np.unique(counts)), counts.min(), counts.max()))
    return _fmt(cre), _fmt(cim), k

ABOUT = ("Deep views fixed by the seeded zoom walk of tests/golden/make_views.py: from the full view "
         "(centre -0.5+0i, width 4) render res x res, move to a uniformly random escaped sample with "
         "count >= 0.9 N (splitmix64(seed)); rule 'boundary' prefers samples next to a non-escaped one, then "
         "band edges; rule 'detail' prefers the samples whose 5 x 5 neighbourhood holds the most distinct "
         "counts; divide the width by 10 until the next width is below the target; use the target width. "
         "A depth with no escaped sample (detail: or one count only) backs the walk off one step to draw "
         "another candidate.")


def main(names):
    path = os.path.join(HERE, "views.json")
    views = json.load(open(path)) if os.path.exists(path) else {}
    views["_about"] = ABOUT
    for name in names:
        seed, N, width, res, rule = SPECS[name]
        if rule == "detail":
            re, im, steps = detail_walk(seed, N, width, res)
        else:
            re, im, steps = O.zoom_walk(seed, N, width, res)
        # re-read before writing: several generator processes may share the file
        views = json.load(open(path)) if os.path.exists(path) else {"_about": ABOUT}
        views[name] = {"seed": seed, "N": N, "target_width": width, "width": width, "res": res, "steps": steps,
                       "rule": rule, "re": re, "im": im}
        print(name, steps, re[:40], im[:40], flush=True)
        with open(path, "w") as f:
            json.dump(views, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:] or list(SPECS))
