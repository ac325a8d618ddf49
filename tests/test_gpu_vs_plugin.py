"""The bm3d VapourSynth plugin on the GPU, driven by the test host as a
VapourSynth core would drive it (plugin init, function invocation from
argument maps, arInitial/arAllFramesReady frame requests, host-memory
frames, frame properties), bit-exact against the oracle:

* Gray float Basic -> Final (profile np);
* RGB48 through RGB2OPP -> Basic(np, matrix=100) -> Final -> OPP2RGB, with
  the BM3D_OPP / _Matrix properties each stage sets;
* VBasic -> VAggregate (radius 1): every stack and every aggregated frame;
* frames requested from several threads at once (fmParallel);
* the BM3D_OPP warning reaching the core's log as mtWarning."""
import numpy as np
import pytest

import oracle as O
import synth
import vshost as V
from oracle_chains import bm3d_frame, make
from test_gpu_vbm3d import _oracle_stack

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def host(gpu):
    h = V.Host()
    h.load()
    return h


def _eq(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    bad = (a.view(np.uint32) != b.view(np.uint32)) if a.dtype == np.float32 else a != b
    assert not bad.any(), f"{what}: {int(bad.sum())} samples differ, first at {np.argwhere(bad)[:3].tolist()}"


def test_gray_basic_final_through_plugin(host, gpu):
    h, w = 56, 72
    frames = [synth.gray_float(h, w, 10.0, seed=4300 + k)[0] for k in range(3)]
    src = host.source(V.GRAY, V.FLOAT, 32, [[f] for f in frames])
    basic = host.invoke("Basic", input=src, profile="np")
    final = host.invoke("Final", input=src, ref=basic, profile="np")
    assert final.info["width"] == w and final.info["height"] == h and final.info["num_frames"] == 3
    fb = make("Basic", gpu.GRAYS, w, h, profile="np")
    ff = make("Final", gpu.GRAYS, w, h, profile="np")
    for n in range(3):
        rb = bm3d_frame(fb, gpu.GRAYS, frames[n][None])
        _eq(basic.get_frame(n)[0][0], rb[0], f"Basic frame {n}")
        _eq(final.get_frame(n)[0][0], bm3d_frame(ff, gpu.GRAYS, frames[n][None], rb)[0], f"Final frame {n}")


def test_rgb48_chain_through_plugin(host, gpu):
    h, w = 48, 64
    rgb, _ = synth.rgb16(h, w, 10.0, seed=4310)
    src = host.source(V.RGB, V.INTEGER, 16, [list(rgb)])
    opp = host.invoke("RGB2OPP", input=src)
    assert opp.info["color_family"] == V.YUV and opp.info["bits"] == 16
    basic = host.invoke("Basic", input=opp, profile="np", sigma=[10.0, 10.0, 10.0], matrix=100)
    final = host.invoke("Final", input=opp, ref=basic, profile="np", sigma=[10.0, 10.0, 10.0], matrix=100)
    out = host.invoke("OPP2RGB", input=final)
    planes, props = opp.get_frame(0)
    want_opp = O.rgb2opp(rgb, 16, 16, True)
    _eq(np.stack(planes), want_opp, "RGB2OPP")
    assert props["BM3D_OPP"] == 1 and props["_Matrix"] == 2
    fmt = gpu.YUV444P16
    fb = make("Basic", fmt, w, h, profile="np", matrix=100)
    ff = make("Final", fmt, w, h, profile="np", matrix=100)
    rb = bm3d_frame(fb, fmt, want_opp)
    rf = bm3d_frame(ff, fmt, want_opp, rb)
    fplanes, fprops = final.get_frame(0)
    _eq(np.stack(fplanes), rf, "Final")
    assert fprops["BM3D_OPP"] == 1
    oplanes, oprops = out.get_frame(0)
    _eq(np.stack(oplanes), O.opp2rgb(rf, 16, 16, True), "OPP2RGB")
    assert "BM3D_OPP" not in oprops and oprops["_Matrix"] == 0


def test_vbasic_vaggregate_through_plugin(host, gpu):
    T, H, W, radius = 4, 40, 48, 1
    noisy, _ = synth.moving_video(T, H, W, 15.0, seed=4320, planes=1)
    src = host.source(V.GRAY, V.FLOAT, 32, [[noisy[t, 0]] for t in range(T)])
    vb = host.invoke("VBasic", input=src, radius=radius)
    agg = host.invoke("VAggregate", input=vb, radius=radius, sample=1)
    assert vb.info["height"] == H * (2 * radius + 1) * 2
    f = make("VBasic", gpu.GRAYS, W, H, radius=radius)
    stacks = []
    for n in range(T):
        planes, props = vb.get_frame(n)
        got = planes[0].reshape(-1, H, W)
        want, s0, s1 = _oracle_stack(f, noisy, noisy, n, radius)
        _eq(got[s0:s1], want[0, s0:s1], f"VBasic frame {n}")
        assert props["BM3D_V_radius"] == radius and props["BM3D_V_process"] == [1, 1, 1]
        stacks.append(got)
    for n in range(T):
        first, last = max(0, n - radius), min(T - 1, n + radius)
        want = O.vaggregate_plane(np.stack(stacks[first:last + 1]), radius, n - first)
        planes, props = agg.get_frame(n)
        _eq(planes[0], want, f"VAggregate frame {n}")
        assert "BM3D_V_radius" not in props


def test_frame_parallel_requests(host, gpu):
    h, w, N = 48, 64, 8
    frames = [synth.gray_float(h, w, 10.0, seed=4330 + k)[0] for k in range(N)]
    src = host.source(V.GRAY, V.FLOAT, 32, [[f] for f in frames])
    par = host.invoke("Final", input=src, ref=host.invoke("Basic", input=src, profile="np"), profile="np")
    par.get_frames_parallel(N, 4)
    seq = host.invoke("Final", input=src, ref=host.invoke("Basic", input=src, profile="np"), profile="np")
    for n in range(N):
        _eq(par.get_frame(n)[0][0], seq.get_frame(n)[0][0], f"frame {n}")


def test_opp_warning_reaches_core_log(host, gpu):
    h, w = 32, 40
    rgb, _ = synth.rgb16(h, w, 10.0, seed=4340)
    opp = O.rgb2opp(rgb, 16, 16, True)
    src = host.source(V.YUV, V.INTEGER, 16, [list(opp)], props={"BM3D_OPP": 1})
    before = len(host.log())
    host.invoke("Basic", input=src, profile="fast").get_frame(0)
    new = host.log()[before:]
    assert any(t == 2 and "BM3D_OPP=1" in m for t, m in new), new
