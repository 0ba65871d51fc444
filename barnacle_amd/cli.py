"""Command line renderer: the reference's Program.fs (`-i/--input`,
`-o/--output`, Program.fs:6-27) on the GPU.

    python -m barnacle_amd -i scene.json -o image.png

Scene.Load (Extensions/Scene/Loader.fs:277-281), then Scene.Render(0, output)
(Extensions/Scene/Render.fs:10-19): the integrator named by the scene's
"integrator.type" renders the film on the GPU (path-tracing, pssmlt, direct,
normal), timed like the reference around the render call only, and the film
is saved through Film.PostProcess (.png) or as linear floats (.pfm).
"""
from __future__ import annotations

import argparse
import sys
import time

from . import film as film_io
from . import native, scene


def render(sc, ctx: native.Context):
    """The integrator of the scene's JSON on the GPU; returns the film
    (Film.Pixels order) and the render time in seconds."""
    ds = ctx.scene(sc)
    kind = sc.integrator["type"]
    t0 = time.perf_counter()
    if kind == "path-tracing":
        img = ds.render_pt(sc.pt_params())
    elif kind == "pssmlt":
        img, _ = ds.render_pssmlt(sc.mlt_params())
    elif kind == "direct":
        img = ds.render_direct(sc.pt_params())
    elif kind == "normal":
        img = ds.render_normal(sc.pt_params())
    else:
        raise scene.SceneError(f"Unknown integrator type: {kind}")
    return img, time.perf_counter() - t0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="barnacle_amd", description="GPU renderer for Barnacle scene files")
    ap.add_argument("-i", "--input", required=True, help="json file containing scene information")
    ap.add_argument("-o", "--output", required=True, help="output filename (.png tone-mapped, .pfm linear)")
    ap.add_argument("--pfm", default=None, help="also write the linear film to this .pfm file")
    ap.add_argument("--device", type=int, default=0, help="HIP device")
    args = ap.parse_args(argv)
    sc = scene.load_scene(args.input)
    print(f"Loaded scene from {args.input}", flush=True)
    ctx = native.Context(args.device)
    try:
        img, secs = render(sc, ctx)
    finally:
        ctx.close()
    print(f"Render time: {secs:f} seconds", flush=True)
    w, h, tone = sc.film
    film_io.save(args.output, img, w, h, tone)
    if args.pfm:
        film_io.write_pfm(args.pfm, img, w, h)
    return 0


if __name__ == "__main__":
    sys.exit(main())
