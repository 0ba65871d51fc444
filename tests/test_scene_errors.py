"""Scene-loading failures carry the reference's messages (Loader.fs:73-237,
failwith), and the oracle refuses a scene with no light the way
LightSamplerBase does (Base/LightSampler.fs:8-9)."""
import copy

import pytest

from barnacle_amd import configs, scene


def base():
    return configs.cornell_box(16, 16, 1, 2)


@pytest.mark.parametrize("path,value,message", [
    (("materials", 0, "type"), "velvet", "Unknown material type: velvet"),
    (("lights", 0, "type"), "spot", "Unknown light type: spot"),
    (("primitives", 0, "type"), "torus", "Unknown primitive type: torus"),
    (("integrator", "type"), "bdpt", "Unknown integrator type: bdpt"),
    (("camera", "type"), "fisheye", "Unknown camera type: fisheye"),
    (("film", "tone-mapping"), "reinhard", "Unknown tone mapping type: reinhard"),
])
def test_unknown_types_fail_with_reference_message(path, value, message):
    info = copy.deepcopy(base())
    d = info
    for k in path[:-1]:
        d = d[k]
    d[path[-1]] = value
    with pytest.raises(scene.SceneError, match=message):
        scene.load_scene(info)


def test_invalid_mesh_and_mutation_strategy():
    info = copy.deepcopy(base())
    info["primitives"][0] = {"type": "mesh"}
    with pytest.raises(scene.SceneError, match="Invalid mesh primitive"):
        scene.load_scene(info)
    info = configs.caustic_scene(16, 16, 1, 2, freq=2, n_bootstrap=64, n_chains=8)
    info["integrator"]["mutation-strategy"] = "Random"
    with pytest.raises(scene.SceneError, match="Unknown mutation strategy: Random"):
        scene.load_scene(info)


def no_light_scene():
    info = copy.deepcopy(base())
    info["instances"] = [{"primitive": i["primitive"], "material": 0} if "light" in i else i
                         for i in info["instances"]]
    return scene.load_scene(info)


def test_oracle_refuses_scene_without_light(orc):
    sc = no_light_scene()
    # Scene.Render builds the light sampler before any integrator runs
    # (Render.fs:13-14): the path tracer and the normal view both fail
    for integrator in (2, 0):
        with pytest.raises(RuntimeError, match="No light primitives found"):
            orc.scene(sc).render(sc.pt_params(), integrator=integrator)
