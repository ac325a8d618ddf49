"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point include/bm3d_amd.h declares, and argument validation of
bm3d_create reproduces the reference's error messages
(source/BM3D_Base.cpp:32-259, source/BM3D_Basic.cpp:44-53,
source/VBM3D_Base.cpp:104-379, source/VAggregate.cpp:30-90,
include/RGB2OPP.h:54-94, include/OPP2RGB.h:54-94) formatted as
VSData::setError does ("bm3d.<Function>: <message>", include/Helper.h:519-523)."""
import ctypes as C
import os
import re
import textwrap

import pytest

import bm3d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = ""
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text += open(os.path.join(ROOT, "include", h)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bm3d_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_declared_symbols():
    L = bm3d.load_library()
    names = _declared_functions()
    assert "bm3d_create" in names and "bm3d_get_frame" in names and "bm3d_lbm_train_step" in names
    for n in names:
        assert hasattr(L, n), n
    assert set(bm3d.EXPORTS) <= set(names)


def _create(name, info, ref=None, **kw):
    L = bm3d.load_library()
    keep = []
    args, nargs = bm3d._make_args(kw, keep)
    h = C.c_void_p()
    err = C.create_string_buffer(1024)
    ri = C.byref(ref) if ref is not None else None
    rc = L.bm3d_create(name.encode(), C.byref(info), ri, args, nargs, C.byref(h), err, 1024)
    return rc, err.value.decode()


def _info(cf=bm3d.YUV, st=bm3d.INTEGER, bits=16, w=64, h=48, n=5, ssw=0, ssh=0):
    return bm3d.ClipInfo(cf, st, bits, ssw, ssh, w, h, n)


E_ARG, E_STATE = 1, 3

CASES = [
    ("Basic", {}, dict(profile="slow"),
     'bm3d.Basic: Unrecognized "profile" specified, should be "fast", "lc", "np", "high" or "vn"'),
    ("Basic", {}, dict(sigma=[10, -1]), 'bm3d.Basic: Invalid "sigma" assigned, must be a non-negative floating point number'),
    ("Basic", {}, dict(block_size=65), 'bm3d.Basic: Invalid "block_size" assigned, must be an integer in [1, 64]'),
    ("Basic", dict(w=16, h=12), dict(block_size=13),
     'bm3d.Basic: Invalid "block_size" assigned, must not exceed width or height of the frame'),
    ("Basic", {}, dict(block_step=9), 'bm3d.Basic: Invalid "block_step" assigned, must be an integer in [1, block_size]'),
    ("Basic", {}, dict(group_size=257), 'bm3d.Basic: Invalid "group_size" assigned, must be an integer in [1, 256]'),
    ("Basic", {}, dict(bm_range=0), 'bm3d.Basic: Invalid "bm_range" assigned, must be a positive integer'),
    ("Basic", {}, dict(bm_step=10), 'bm3d.Basic: Invalid "bm_step" assigned, must be an integer in [1, bm_range]'),
    ("Basic", {}, dict(th_mse=0.0), 'bm3d.Basic: Invalid "th_mse" assigned, must be a positive floating point number'),
    ("Basic", {}, dict(matrix=3), 'bm3d.Basic: Unsupported "matrix" specified'),
    ("Basic", {}, dict(hard_thr=0.0), 'bm3d.Basic: Invalid "hard_thr" assigned, must be a positive floating point number'),
    ("Basic", dict(ssw=1), {},
     "bm3d.Basic: input clip: sub-sampled format is not supported when chroma is processed, convert it to YUV444 "
     "or RGB first. For the best quality, RGB colorspace is recommended as input."),
    ("Basic", dict(bits=20), {},
     "bm3d.Basic: Invalid input clip, only 8-16 bit integer or 32 bit float formats supported"),
    ("VBasic", {}, dict(radius=17), 'bm3d.VBasic: Invalid "radius" assigned, must be an integer in [1, 16]'),
    ("VBasic", {}, dict(ps_num=9), 'bm3d.VBasic: Invalid "ps_num" assigned, must be an integer in [1, group_size]'),
    ("VBasic", {}, dict(ps_range=0), 'bm3d.VBasic: Invalid "ps_range" assigned, must be a positive integer'),
    ("VBasic", {}, dict(ps_step=5), 'bm3d.VBasic: Invalid "ps_step" assigned, must be an integer in [1, ps_range]'),
    ("VAggregate", {}, {}, "bm3d.VAggregate: Invalid input clip, only accept 32 bit float format clip from "
                          "bm3d.VBasic or bm3d.VFinal"),
    ("VAggregate", dict(st=bm3d.FLOAT, bits=32), dict(radius=0),
     'bm3d.VAggregate: Invalid "radius" assigned, must be an integer in [1, 16]'),
    ("VAggregate", dict(st=bm3d.FLOAT, bits=32, cf=bm3d.RGB), {},
     "bm3d.VAggregate: Invalid input clip, must be of Gray or YUV color family"),
    ("RGB2OPP", {}, {}, "bm3d.RGB2OPP: Invalid input clip, must be of RGB color family"),
    ("RGB2OPP", dict(cf=bm3d.RGB), dict(sample=2),
     "bm3d.RGB2OPP: Invalid 'sample' assigned, must be 0 (integer sample type) or 1 (float sample type)"),
    ("OPP2RGB", dict(cf=bm3d.RGB), {}, "bm3d.OPP2RGB: Invalid input clip, must be of YUV color family"),
    # no counterpart in the reference: the arithmetic selector and the size
    # limit of the packed (y << 16 | x) match positions
    ("Basic", {}, dict(arith=2), 'bm3d.Basic: Invalid "arith" assigned, must be 0 (x86 SSE2 arithmetic) or 1 '
                                 '(scalar arithmetic)'),
    ("VFinal", {}, dict(arith=-1), 'bm3d.VFinal: Invalid "arith" assigned, must be 0 (x86 SSE2 arithmetic) or 1 '
                                   '(scalar arithmetic)'),
    ("Basic", dict(h=32768), {}, "bm3d.Basic: frames wider than 65535 or taller than 32767 pixels are not supported"),
    ("Final", dict(w=65536), {}, "bm3d.Final: frames wider than 65535 or taller than 32767 pixels are not supported"),
]


@pytest.mark.parametrize("name,info_kw,args,msg", CASES)
def test_argument_errors(name, info_kw, args, msg):
    ref = _info(**info_kw) if name in ("Final", "VFinal") else None
    rc, err = _create(name, _info(**info_kw), ref, **args)
    assert rc == E_ARG
    assert err == msg


def test_ref_clip_checks():
    rc, err = _create("Basic", _info(), _info(w=32))
    assert rc == E_ARG and err == 'bm3d.Basic: input clip and clip "ref" must be of the same width and height'
    rc, err = _create("Basic", _info(), _info(bits=8))
    assert rc == E_ARG and err == 'bm3d.Basic: input clip and clip "ref" must be of the same format'
    rc, err = _create("Final", _info(), _info(n=4))
    assert rc == E_ARG and err == 'bm3d.Final: input clip and clip "ref" must have the same number of frames'


@pytest.mark.parametrize("name,kw", [
    ("Basic", dict(profile="np", sigma=[10, 0, 3], matrix=100)),
    ("Basic", dict(sigma=[0])),  # sigma 0 on YUV disables every plane (allowed)
    ("VBasic", dict(radius=3, ps_num=2, ps_range=4)),
    ("RGB2OPP", dict(sample=1)),
])
def test_valid_arguments_pass_validation(name, kw):
    """Valid arguments pass validation; without bm3d_init the call then stops
    at the device step with BM3D_E_STATE (no GPU needed for this test)."""
    info = _info(cf=bm3d.RGB) if name == "RGB2OPP" else _info()
    rc, err = _create(name, info, **kw)
    assert rc == E_STATE, err


def test_host_arith_default_and_forced_probe_failure():
    """On this x86 host the rcpps probe tabulates, so the default is the SSE2
    arithmetic (0).  With BM3D_AMD_HOST_RCP=none the probe reports failure:
    the default becomes the scalar arithmetic (1), arith=1 is accepted and
    arith=0 is refused (checked in a fresh process: the probe runs once)."""
    import subprocess
    import sys
    import platform
    if platform.machine() in ("x86_64", "AMD64"):
        assert bm3d.host_arith() == 0
    code = textwrap.dedent("""
        import sys
        sys.path.insert(0, %r)
        sys.path.insert(0, %r)
        import bm3d, test_host_abi as T
        print(bm3d.host_arith())
        print(T._create("Final", T._info(), T._info(), arith=0))
        print(T._create("Final", T._info(), T._info(), arith=1))
    """) % (os.path.join(ROOT, "vapoursynth-bm3d_amd"), os.path.join(ROOT, "tests"))
    env = dict(os.environ, BM3D_AMD_HOST_RCP="none")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    assert lines[0] == "1"
    assert lines[1] == str((E_ARG, 'bm3d.Final: "arith=0" needs the host\'s SSE2 rcpps, which could not be '
                                   'tabulated on this host'))
    assert lines[2].startswith(f"({E_STATE},")  # valid: stops at the device step without bm3d_init
