"""The VapourSynth plugin shim (vapoursynth-bm3d_amd/_build/libbm3d_vs.so)
loaded by the test host the way a VapourSynth core loads plugins (dlopen +
VapourSynthPluginInit2), no GPU needed:

* identity and registrations equal the reference's
  (source/VSPlugin.cpp:424-517): identifier, namespace, description,
  version 10.1, API v4, and the seven functions with their argument strings
  and return type, in order;
* the core's own checks of those strings (required ref for Final/VFinal,
  unknown argument names, arrays, int-to-float conversion) reach the plugin
  as the reference's would;
* the plugin reports the reference's argument errors through mapSetError
  ("bm3d.<Function>: <message>") before any device work."""
import numpy as np
import pytest

import vshost as V

REFERENCE_FUNCTIONS = [
    ("RGB2OPP", "input:vnode;sample:int:opt;"),
    ("OPP2RGB", "input:vnode;sample:int:opt;"),
    ("Basic", "input:vnode;ref:vnode:opt;profile:data:opt;sigma:float[]:opt;block_size:int:opt;block_step:int:opt;"
              "group_size:int:opt;bm_range:int:opt;bm_step:int:opt;th_mse:float:opt;hard_thr:float:opt;"
              "matrix:int:opt;"),
    ("Final", "input:vnode;ref:vnode;profile:data:opt;sigma:float[]:opt;block_size:int:opt;block_step:int:opt;"
              "group_size:int:opt;bm_range:int:opt;bm_step:int:opt;th_mse:float:opt;matrix:int:opt;"),
    ("VBasic", "input:vnode;ref:vnode:opt;profile:data:opt;sigma:float[]:opt;radius:int:opt;block_size:int:opt;"
               "block_step:int:opt;group_size:int:opt;bm_range:int:opt;bm_step:int:opt;ps_num:int:opt;"
               "ps_range:int:opt;ps_step:int:opt;th_mse:float:opt;hard_thr:float:opt;matrix:int:opt;"),
    ("VFinal", "input:vnode;ref:vnode;profile:data:opt;sigma:float[]:opt;radius:int:opt;block_size:int:opt;"
               "block_step:int:opt;group_size:int:opt;bm_range:int:opt;bm_step:int:opt;ps_num:int:opt;"
               "ps_range:int:opt;ps_step:int:opt;th_mse:float:opt;matrix:int:opt;"),
    ("VAggregate", "input:vnode;radius:int:opt;sample:int:opt;"),
]


@pytest.fixture(scope="module")
def host():
    h = V.Host()
    h.load()
    return h


def _gray(host, h=24, w=32, n=2):
    rng = np.random.RandomState(5)
    return host.source(V.GRAY, V.FLOAT, 32, [[rng.rand(h, w).astype(np.float32)] for _ in range(n)])


def test_plugin_identity(host):
    info = host.plugin_info()
    assert info["identifier"] == "com.vapoursynth.bm3d"
    assert info["namespace"] == "bm3d"
    assert info["name"] == "Implementation of BM3D denoising filter for VapourSynth."
    assert info["version"] == (10 << 16) | 1
    assert info["api_version"] >> 16 == 4
    assert info["flags"] == 0


def test_registered_functions_equal_the_reference(host):
    fns = host.functions()
    assert [(n, a) for n, a, _ in fns] == REFERENCE_FUNCTIONS
    assert all(r == "clip:vnode;" for _, _, r in fns)


@pytest.mark.parametrize("fn,kw,msg", [
    ("Final", {}, "bm3d.Final: argument ref is required"),
    ("VFinal", {}, "bm3d.VFinal: argument ref is required"),
    ("Basic", dict(arith=1), "bm3d.Basic: Function does not take argument(s) named arith"),
    ("Basic", dict(block_size=[8, 8]), "bm3d.Basic: argument block_size is not of array type but more than one "
                                       "value was supplied"),
    ("Basic", dict(profile=3), "bm3d.Basic: argument profile is not of the correct type"),
])
def test_core_checks_of_the_registered_strings(host, fn, kw, msg):
    with pytest.raises(V.HostError) as e:
        host.invoke(fn, input=_gray(host), **kw)
    assert str(e.value) == msg


@pytest.mark.parametrize("fn,kw,msg", [
    ("Basic", dict(profile="slow"),
     'bm3d.Basic: Unrecognized "profile" specified, should be "fast", "lc", "np", "high" or "vn"'),
    ("Basic", dict(sigma=[10.0, -1.0]), 'bm3d.Basic: Invalid "sigma" assigned, must be a non-negative floating point number'),
    ("Basic", dict(sigma=[10, -1]), 'bm3d.Basic: Invalid "sigma" assigned, must be a non-negative floating point number'),
    ("Basic", dict(block_size=65), 'bm3d.Basic: Invalid "block_size" assigned, must be an integer in [1, 64]'),
    ("Basic", dict(th_mse=0.0), 'bm3d.Basic: Invalid "th_mse" assigned, must be a positive floating point number'),
    ("VBasic", dict(radius=17), 'bm3d.VBasic: Invalid "radius" assigned, must be an integer in [1, 16]'),
    ("RGB2OPP", {}, "bm3d.RGB2OPP: Invalid input clip, must be of RGB color family"),
    ("VAggregate", dict(radius=0), 'bm3d.VAggregate: Invalid "radius" assigned, must be an integer in [1, 16]'),
])
def test_plugin_argument_errors(host, fn, kw, msg):
    with pytest.raises(V.HostError) as e:
        host.invoke(fn, input=_gray(host), **kw)
    assert str(e.value) == msg


def test_ref_clip_mismatch(host):
    with pytest.raises(V.HostError) as e:
        host.invoke("Final", input=_gray(host), ref=_gray(host, w=40))
    assert str(e.value) == 'bm3d.Final: input clip and clip "ref" must be of the same width and height'
