"""Launch parameters (persistent grid, lane refill, LDS stack window, slots in
flight) change the schedule, never the image: films stay bit-identical, and
the autotuner leaves a valid setting on the context."""
import numpy as np
import pytest

from barnacle_amd import configs, scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mesh_job(ctx):
    sc = scene.load_scene(configs.config1(width=48, height=40, spp=3, max_depth=16, freq=16))
    return ctx.scene(sc), sc.pt_params()


SETTINGS = [
    {"stack_k": 32},
    {"stack_k": 32, "trace_blocks_per_cu": 2, "refill": 16},
    {"stack_k": 16, "trace_blocks_per_cu": 5, "refill": 64, "chunk_slots": 1000},
    {"material_sort": 1},
    {"material_sort": 0},
]


@pytest.mark.parametrize("opts", SETTINGS)
def test_launch_parameters_do_not_change_film(ctx, mesh_job, opts):
    ds, p = mesh_job
    ref = ds.render_pt(p)
    try:
        for k, v in opts.items():
            ctx.set_option(k, v)
        got = ds.render_pt(p)
        films, _ = ctx.render_batch([(ds, p), (ds, p)])
    finally:
        for k, v in {"stack_k": 16, "trace_blocks_per_cu": 7, "refill": 40, "chunk_slots": 64 << 20,
                     "material_sort": -1}.items():
            ctx.set_option(k, v)
    assert np.array_equal(got.view(np.int32), ref.view(np.int32))
    for f in films:
        assert np.array_equal(f.view(np.int32), ref.view(np.int32))


def test_autotune(ctx, mesh_job):
    ds, p = mesh_job
    ref = ds.render_pt(p)
    try:
        r = ctx.autotune([(ds, p)])
        assert r["trace_blocks_per_cu"] in (4, 5, 6, 7, 8)
        assert r["refill"] in (24, 32, 40, 48)
        assert r["stack_k"] in (16, 32)
        assert r["trials"] >= 8
        assert 0 < r["best_ms"] <= r["baseline_ms"]
        assert np.array_equal(ds.render_pt(p).view(np.int32), ref.view(np.int32))
        rb = ctx.autotune([(ds, p), (ds, p)], batched=True)
        assert rb["trials"] >= 8
    finally:
        for k, v in {"stack_k": 16, "trace_blocks_per_cu": 7, "refill": 40, "chunk_slots": 64 << 20,
                     "material_sort": -1}.items():
            ctx.set_option(k, v)
