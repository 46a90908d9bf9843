"""LBVH node logic and the stable radix ranking, checked on the CPU."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_hierarchy_and_ropes_host_check(tmp_path):
    exe = str(tmp_path / "bvh_host_check")
    src = os.path.join(ROOT, "tests", "bvh_host_check.hip")
    cmd = ["/opt/rocm/bin/hipcc", "-std=c++17", "-O1", "--offload-arch=gfx950",
           "-I", os.path.join(ROOT, "distance3d_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
           src, "-o", exe]
    try:
        subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    except FileNotFoundError:
        pytest.skip("hipcc not available")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().endswith("OK")


def _emulate_stable_scatter(keys, shift, tile, block=256, wave=64):
    """The ranking scheme of rs_scatter_kernel, step for step."""
    n = len(keys)
    digits = (keys >> shift) & 255
    tiles = (n + tile - 1) // tile
    hist = np.zeros((256, tiles), dtype=np.int64)
    for b in range(tiles):
        d = digits[b * tile:(b + 1) * tile]
        hist[:, b] = np.bincount(d, minlength=256)
    flat = hist.reshape(-1)
    offs = (np.cumsum(flat) - flat).reshape(256, tiles)
    out = np.empty(n, dtype=keys.dtype)
    for b in range(tiles):
        run = offs[:, b].copy()
        for j in range(tile // block):
            base = b * tile + j * block
            idx = np.arange(base, min(base + block, n))
            if len(idx) == 0:
                break
            d = digits[idx]
            wcnt = np.zeros((block // wave, 256), dtype=np.int64)
            pos = np.empty(len(idx), dtype=np.int64)
            for w in range(block // wave):
                sl = slice(w * wave, min((w + 1) * wave, len(idx)))
                dw = d[sl]
                for q, dig in enumerate(dw):
                    rank = np.sum(dw[:q] == dig)
                    pos[w * wave + q] = run[dig] + wcnt[:w, dig].sum() + rank
                if len(dw):
                    wcnt[w] = np.bincount(dw, minlength=256)
            out[pos] = keys[idx]
            run += wcnt.sum(axis=0)
    return out


def test_stable_ranking_scheme_sorts():
    rs = np.random.RandomState(3)
    keys = rs.randint(0, 1 << 30, 3000).astype(np.uint32)
    keys[::7] = keys[0]
    k = keys.copy()
    for p in range(4):
        k = _emulate_stable_scatter(k, 8 * p, tile=1024)
    assert np.array_equal(k, np.sort(keys))
