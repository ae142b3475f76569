"""The C ABI as a C program sees it (tests/abi_check.c, compiled against
include/*.h with gcc): on the CPU it must compile, link and run the host
set-up; on the GPU its whole run must equal the same run made through the
Python binding, value for value."""
import os
import subprocess

import numpy as np
import pytest

import gragles_3d_amd as G
from gragles_3d_amd import build as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "abi_check.c")
BIN = os.path.join(ROOT, "tests", "_bin", "abi_check")
ARGS = (27, 12, 3, 11, 4)  # grains, points per grain, border, seed, steps


def _parse(text):
    out = {"grain": {}}
    for line in text.splitlines():
        k, _, v = line.partition(" ")
        if k == "grain":
            g, st, vol, area = v.split()
            out["grain"][int(g)] = (int(st), float(vol), float(area))
        else:
            out[k] = v
    return out


def _host_expect(p, seed):
    idf, bounds, exists = G.generate_voronoi(p, seed)
    offs, ids = G.initial_neighbour_lists(p, bounds, exists)
    return idf, bounds, exists, offs, ids


def test_abi_program_compiles_and_runs_host_setup():
    # compile here too: the CPU suite checks that the headers build as C and link
    B.build_abi_check(SRC, BIN)
    r = subprocess.run([BIN] + [str(a) for a in ARGS] + ["0"], capture_output=True, text=True,
                       env=dict(os.environ, HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", "")))
    out = _parse(r.stdout)
    assert r.returncode in (0, 77), r.stderr
    p = G.derive_params(*ARGS[:3])
    assert int(out["ngridpoints"]) == p.ngridpoints
    assert float(out["dt"]) == p.dt and float(out["h"]) == p.h and float(out["delta"]) == p.delta
    idf, _, _, offs, ids = _host_expect(p, ARGS[3])
    assert int(out["idfield_sum"]) == int(idf.astype(np.int64).sum())
    assert int(out["list_entries"]) == len(ids)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [G.FP64_PLAIN, G.FP32_BATCHED])
def test_abi_program_equals_python_binding(precision):
    print("abi program n=%d ppg=%d border=%d seed=%d steps=%d precision=%d" % (ARGS + (precision,)), flush=True)
    assert os.path.exists(BIN), "tests/_bin/abi_check is not built: run __graft_entry__.build()"
    r = subprocess.run([BIN] + [str(a) for a in ARGS] + [str(precision)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    out = _parse(r.stdout)
    ctx, p, inp = G.setup_simulation(*ARGS[:4], precision=precision)
    ctx.timestep(ARGS[4], use_graph=True)
    st, vol, area = ctx.stats()
    hist = ctx.count_history()
    assert out["count_history"].split() == [str(c) for c in hist]
    for g in range(1, ARGS[0] + 1):
        assert out["grain"][g] == (int(st[g]), float(vol[g]), float(area[g])), g
    g, *rest = out["phi_grain"].split()
    b, phi = ctx.phi(int(g), 1)
    assert [int(x) for x in rest[:6]] == b.tolist() and int(rest[6]) == phi.size
    s = 0.0
    for v in phi:  # the program's left-to-right sum
        s += float(v)
    assert float(rest[7]) == s
    assert int(out["redistancing_again"]) == G.levelset.GG_ERR_STATE
    ctx.close()
