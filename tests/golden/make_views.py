"""Regenerate tests/golden/views.json: the deep views fixed by the seeded zoom walk.

The walk (oracle/oracle.c, orc_zoom_walk) starts from the reference's full view
(centre -0.5+0i, width 4, mandelbrot.cpp:13-14), renders res x res at the
current width, and moves the centre to a uniformly random escaped sample with
count >= 0.9 N (splitmix64(seed)), preferring samples next to a non-escaped sample
(the set's boundary), then samples on a band edge; the width is divided by 10
until the next width would fall below the target, which is then used as is.  A
depth whose render has no escaped sample backs the walk off one step, where another
candidate is drawn from the same splitmix64 stream.

Usage: python tests/golden/make_views.py [name ...]   (default: all)
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402

# name: (seed, N, target width, walk resolution)
SPECS = {
    "t11": (11, 3000, "1e-16", 128),
    "t13": (13, 8000, "1e-25", 128),
    "c2_seed1": (1, 20000, "1e-30", 256),
    "c3_seed1": (1, 50000, "1e-200", 256),
    "c4_seed1": (1, 60000, "1e-40", 256),
    "c5_seed1": (1, 100000, "1e-100", 256),
}

ABOUT = ("Deep views fixed by the seeded zoom walk of tests/golden/make_views.py: from the full view "
         "(centre -0.5+0i, width 4) render res x res, move to a uniformly random escaped sample with "
         "count >= 0.9 N (splitmix64(seed)), preferring samples next to a non-escaped one, then band edges; "
         "divide the width by 10 until the next width is below the target; use the target width. A depth "
         "with no escaped sample backs the walk off one step to draw another candidate.")


def main(names):
    path = os.path.join(HERE, "views.json")
    views = json.load(open(path)) if os.path.exists(path) else {}
    views["_about"] = ABOUT
    for name in names:
        seed, N, width, res = SPECS[name]
        re, im, steps = O.zoom_walk(seed, N, width, res)
        views[name] = {"seed": seed, "N": N, "target_width": width, "width": width, "res": res, "steps": steps,
                       "re": re, "im": im}
        print(name, steps, re[:40], im[:40], flush=True)
        with open(path, "w") as f:
            json.dump(views, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:] or list(SPECS))
