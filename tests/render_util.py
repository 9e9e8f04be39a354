"""Run a RenderScene on the GPU path tracer and on the oracle integrator."""
import numpy as np

import alpine_amd as aa


def gpu_render(sc, spp=None, fp64=False, builder=aa.BUILDER_GPU, acc=None, reset=True):
    if acc is None:
        acc = aa.GpuAccelerator(0)
        acc.append_mesh(sc.verts, sc.tris)
        acc.update_scene(builder=builder)
        acc.set_materials([aa.material(*m) for m in sc.materials], sc.tri_mat)
        acc.set_lights([aa.light(**l) for l in sc.lights], sc.light_sampler)
        acc.set_camera(aa.camera(**sc.camera))
    if reset:
        acc.render_reset(sc.width, sc.height, sc.seed)
    acc.render(spp or sc.spp, sc.max_depth, sc.rr_depth, sc.background, fp64=fp64)
    rgb, r64, total = acc.render_read(fp64=fp64)
    return acc, rgb, r64, total


def oracle_render(sc, spp=None, nthreads=8, states=None, accum=None):
    import oracle
    mats = [oracle.make_material(*m) for m in sc.materials]
    lights = [oracle.make_light(**l) for l in sc.lights]
    cam = oracle.make_camera(**sc.camera)
    os_ = oracle.OracleScene(sc.verts, sc.tris, sc.tri_mat, mats, lights, sc.light_sampler, cam,
                             sc.background)
    if states is None:
        states = oracle.sampler_states(sc.width * sc.height, sc.seed)
    states, accum = os_.render(sc.width, sc.height, spp or sc.spp, sc.max_depth, states, accum,
                               sc.rr_depth, nthreads)
    return states, accum
