"""The C-ABI library loads and exports every symbol include/svpf.h declares, and each
VapourSynth plugin library exports exactly what include/svpflow_vs.h declares (no GPU
calls)."""
import ctypes
import os
import re
import subprocess

import svpf

HERE = os.path.dirname(os.path.abspath(__file__))


def test_library_exports_declared_symbols():
    names = svpf.exported_symbols()
    assert len(names) >= 10
    lib = ctypes.CDLL(svpf.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_plugin_libraries_export_declared_symbols():
    text = open(os.path.join(HERE, "..", "include", "svpflow_vs.h")).read()
    declared = sorted(set(re.findall(r"\b(VapourSynthPluginInit|svpGetVersion)\s*\(", text)))
    assert declared == ["VapourSynthPluginInit", "svpGetVersion"]
    for name in ("libsvpflow1_vs.so", "libsvpflow2_vs.so"):
        path = os.path.join(HERE, "..", "open-svpflow_amd", name)
        out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
        exported = sorted(line.split()[-1] for line in out.splitlines() if line.split()[1] in "TW")
        assert exported == declared, f"{name} exports {exported}"


def test_option_parsing_matches_oracle(oracle):
    # product-side SuperOpts::from_opt vs the oracle's restatement (host logic, no GPU)
    for opt in (None, "{pel:1}", "{pel:4, gpu:0}", "{gpu:2}", "{pel:2, full:1}", "{pel:2, rc:true}",
                "{ /* c */ pel : 1 // x\n }", "{pel:2.7}", "{pel:-}", "garbage", "{pel:2, pel:1}"):
        for w, h in ((1920, 1080), (640, 360), (854, 480), (34, 18)):
            o, err = oracle.super_opts(opt, w, h)
            try:
                p, sdata = svpf.super_params(opt, w, h)
            except svpf.SvpfError as e:
                assert o is None and str(e) == err
                continue
            assert o is not None, opt
            assert (p.width, p.height, p.pel, p.gpu, p.scale_down, p.scale_up, p.full, p.rc, p.levels) == \
                tuple(o[k] for k in ("width", "height", "pel", "gpu", "scale_down", "scale_up", "full", "rc",
                                     "levels"))
            assert sdata == o["sdata"]
            assert svpf.super_height(p) == o["super_h"]
    for opt in ("{pel:3}", "{gpu:3}", "{full:false}"):
        o, err = oracle.super_opts(opt, 64, 64)
        try:
            svpf.super_params(opt, 64, 64)
            assert False, "expected an error"
        except svpf.SvpfError as e:
            assert str(e) == err
