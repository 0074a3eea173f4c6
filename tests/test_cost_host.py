"""CPU parity of the HIP search's whole-candidate cost (analyse_cost.h, built for the
host) against the oracle's candidate_cost on random candidates: every path of the
reference's selection -- interior in-level blocks, satd_luma_clamp and the SAD row
breaks at the level borders, the edge-shifted path for out-of-range vectors, chroma
interior/non-square/fallback-128 paths -- at pel 1 and 2, levels 0-2, SAD and SATD.
The GPU runs the same code for candidates near the borders (its split fast path for
interior candidates is covered by the GPU Analyse tests)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import synth

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")


@pytest.fixture(scope="module")
def host_lib():
    os.makedirs(BUILD, exist_ok=True)
    so = os.path.join(BUILD, "libcosthost.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", so,
                    os.path.join(HERE, "hostcost", "cost_host.cpp")], check=True)
    L = C.CDLL(so)
    u8p = C.POINTER(C.c_uint8)
    L.svpf_host_cand_cost.argtypes = [C.c_int] * 7 + [u8p, C.c_int, C.c_int] * 6 + [C.c_int] * 7 + \
        [C.POINTER(C.c_uint32)] * 2
    return L


def reduce_dim(d, n):
    for _ in range(n):
        d = 2 * (d // 4)
    return d


def level_offsets(h, pel, level):
    y = cy = 0
    for lv in range(level):
        hh = reduce_dim(h, lv)
        sub = pel * pel if lv == 0 else 1
        y += hh * sub
        cy += max(hh // 2, 1) * sub
    return y, cy


@pytest.mark.parametrize("pel", [1, 2])
def test_cost_paths_match_oracle(oracle, host_lib, pel):
    w, h = 96, 64
    frames = synth.clip(w, h, 2, seed=77 + pel, vmax=4.0, nobj=(1, 1))
    sopt = f"{{pel:{pel}}}"
    cur = [np.ascontiguousarray(p) for p in oracle.super_build(sopt, frames[0])]
    ref = [np.ascontiguousarray(p) for p in oracle.super_build(sopt, frames[1])]
    so, _ = oracle.super_opts(sopt, w, h)
    levels = so["levels"] if "levels" in so else 3
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint8))
    rng = np.random.default_rng(1234 + pel)
    checked = 0
    for level in range(3):
        lw, lh = reduce_dim(w, level), reduce_dim(h, level)
        y_off, c_off = level_offsets(h, pel, level)
        lpel = pel if level == 0 else 1
        for bw, bh in [(4, 4), (8, 8), (16, 16), (16, 8)]:
            if bw > lw or bh > lh:
                continue
            for _ in range(60):
                px = int(rng.integers(0, lw - bw + 1))
                py = int(rng.integers(0, lh - bh + 1))
                if rng.random() < 0.3:  # the last block of a coarse row/column sits at the level end
                    px = lw - bw
                span = lpel * 24
                mvx = int(rng.integers(-span, span + 1))
                mvy = int(rng.integers(-span, span + 1))
                satd = bool(rng.random() < 0.5)
                chroma = bool(rng.random() < 0.8)
                want = oracle.candidate_cost(w, h, pel, levels, cur, ref, level, px, py, bw, bh, mvx, mvy, satd, chroma)
                l, c = C.c_uint32(), C.c_uint32()
                args = []
                for planes in (cur, ref):
                    for p in planes:
                        args += [ptr(p), p.shape[1], p.size]
                host_lib.svpf_host_cand_cost(level, lw, lh, y_off, c_off, lpel, int(satd), *args, px, py, bw, bh, mvx,
                                             mvy, int(chroma), C.byref(l), C.byref(c))
                assert (l.value, c.value) == want, (
                    f"level {level} block {bw}x{bh} at ({px},{py}) mv ({mvx},{mvy}) satd {satd} chroma {chroma}: "
                    f"got {(l.value, c.value)} want {want}")
                checked += 1
    assert checked > 500
