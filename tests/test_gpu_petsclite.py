"""The PETSc-compatible surface (petsclite + RHSFunction) driven by C programs.

tests/drivers/plane_wave_3d.c follows the reference plane-wave driver
(3d/weno4/plane_wave/hype.c) call for call; its errors must reproduce the
published convergence table (results/main.tex:37-42, three significant
digits, at final time 1) through the fused GPU stepper.
tests/drivers/acoustic_ts.c checks the fused and the generic TSSSP paths
against each other and against the oracle.
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "tests", "drivers")


def run(args, timeout=300):
    print("[case]", " ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("N", [16, 32, 64, 128])
def test_plane_wave_table_through_petsclite(N):
    rows = json.load(open(os.path.join(ROOT, "tests", "golden", "plane_wave_3d_table.json")))["rows"]
    row = [r for r in rows if r["N"] == N][0]
    out = run([os.path.join(DRV, "plane_wave_3d"), str(N), "1.0"]).split()
    l2, li = float(out[1]), float(out[2])
    print(f"[case] N={N} l2={l2:.4e} ({row['l2_text']}) linf={li:.4e} ({row['linf_text']})", flush=True)
    assert abs(l2 / row["l2"] - 1) < 2e-3
    assert abs(li / row["linf"] - 1) < 3e-3


def test_plane_wave_n16_matches_oracle(orc):
    """Same errors as the oracle's reference-structured RHS to ~1e-10."""
    import math
    out = run([os.path.join(DRV, "plane_wave_3d"), "16", "1.0"]).split()
    N, h = 16, 2 * math.pi / 16
    g = orc.grid3d(N, h, orc.WAVE2, cfl=0.5)
    U = orc.plane_wave_init(N, h)
    orc.ts_solve("orc_rhs3d_cb_faithful", g, U, h * 0.5 / 3.0, 1.0, 0, 4)
    l2, li = orc.plane_wave_error(N, h, 1.0, U)
    assert abs(float(out[1]) / l2 - 1) < 1e-9 and abs(float(out[2]) / li - 1) < 1e-9


@pytest.mark.parametrize("gamma", [(1.0, 0.0), (0.85, 0.85)])
def test_fused_generic_and_oracle_agree(tmp_path, orc, acx, gamma):
    N, steps = 24, 5
    res = {}
    for mode in ("fused", "generic", "rk104"):
        f = str(tmp_path / f"{mode}.bin")
        out = run([os.path.join(DRV, "acoustic_ts"), str(N), str(steps), str(gamma[0]), str(gamma[1]), mode, f])
        assert f"steps {steps}" in out
        res[mode] = np.fromfile(f, dtype=np.float64).reshape(N, N, N, 4)
    d = np.linalg.norm(res["fused"] - res["generic"]) / np.linalg.norm(res["fused"])
    print(f"[case] gamma={gamma} fused vs generic rel {d:.3e}", flush=True)
    assert d < 1e-13
    grid = acx.Grid(n=(N, N, N), h=1.0 / N, gamma_hi=gamma[0], gamma_lo=gamma[1])
    U = acx.field_gaussians(grid, 2025, npulse=4, amp=(1e-4, 1e-3), sigma=(0.02, 0.08))
    dt = 0.4 * (1.0 / N) / 3.0
    g = orc.grid3d(N, 1.0 / N, orc.ACOUSTIC4, gamma)
    ref = U.copy()
    orc.ts_solve("orc_rhs3d_cb_faithful", g, ref, dt, steps * dt, 0, 4)
    e = np.linalg.norm(res["fused"] - ref) / np.linalg.norm(ref)
    print(f"[case] gamma={gamma} fused vs oracle rel {e:.3e}", flush=True)
    assert e < 1e-12
    ref4 = U.copy()
    orc.ts_solve("orc_rhs3d_cb_faithful", g, ref4, dt, steps * dt, 1, 0)
    e4 = np.linalg.norm(res["rk104"] - ref4) / np.linalg.norm(ref4)
    print(f"[case] gamma={gamma} rk104 vs oracle rel {e4:.3e}", flush=True)
    assert e4 < 1e-12


def test_restart_files_and_vtk(tmp_path):
    """Monitor output and restart of the reference drivers
    (3d/weno4/plane_wave/monitor_function.c:12-62, hype.c:144-196): a run
    resumed from restart1.bin (VecLoad) ends bitwise where the uninterrupted
    run ends; the binary files follow the PETSc layout (class id 1211214,
    int32 length, big-endian float64); the ASCII VTK files hold the grid and
    the final state."""
    N, steps, ri, wi = 20, 10, 4, 5
    (tmp_path / "plot").mkdir()
    out = run([os.path.join(DRV, "restart_3d"), str(N), str(steps), str(ri), str(wi), "run", str(tmp_path)])
    assert "10 t = " in out
    run([os.path.join(DRV, "restart_3d"), str(N), str(steps), str(ri), str(wi), "resume", str(tmp_path)])
    a = (tmp_path / "final.bin").read_bytes()
    b = (tmp_path / "final_resumed.bin").read_bytes()
    assert a == b
    n = N ** 3 * 4
    for name in ("final.bin", "restart1.bin", "restart2.bin"):
        raw = (tmp_path / name).read_bytes()
        assert len(raw) == 8 + 8 * n
        head = np.frombuffer(raw[:8], dtype=">i4")
        assert head[0] == 1211214 and head[1] == n
    assert (tmp_path / "final.bin.info").read_text().strip() == "-vecload_block_size 4"
    final = np.frombuffer(a[8:], dtype=">f8").astype(np.float64)
    r1 = np.frombuffer((tmp_path / "restart1.bin").read_bytes()[8:], dtype=">f8")
    r2 = np.frombuffer((tmp_path / "restart2.bin").read_bytes()[8:], dtype=">f8")
    assert np.abs(r2).max() > 0 and not np.array_equal(r1, r2)   # step 0 vs step 8
    # VTK written every 5 steps and at the end
    for s in (0, 5, 10):
        assert (tmp_path / "plot" / f"sol-{s:05d}.vtk").exists()
    lines = (tmp_path / "plot" / f"sol-{steps:08d}.vtk").read_text().splitlines()
    assert lines[0] == "# vtk DataFile Version 2.0" and lines[2] == "ASCII"
    assert lines[3] == "DATASET STRUCTURED_GRID" and lines[4] == f"DIMENSIONS {N} {N} {N}"
    assert lines[5] == f"POINTS {N ** 3} double"
    pts = np.array([[float(t) for t in ln.split()] for ln in lines[6:6 + N ** 3]])
    h = 1.0 / N
    i = np.arange(N ** 3)
    want = np.stack([(i % N + 0.5) * h, ((i // N) % N + 0.5) * h, (i // N ** 2 + 0.5) * h], 1)
    assert np.max(np.abs(pts - want)) < 1e-14
    k = 6 + N ** 3
    assert lines[k] == f"POINT_DATA {N ** 3}" and lines[k + 1] == "SCALARS P_U_V_W double 4"
    assert lines[k + 2] == "LOOKUP_TABLE default"
    vals = np.array([[float(t) for t in ln.split()] for ln in lines[k + 3:k + 3 + N ** 3]]).ravel()
    assert np.array_equal(vals, final)
    # XML structured grid (PetscViewerVTKOpen): header, then appended raw blocks
    # (UInt64 byte count + little-endian Float64 data) at the declared offsets
    raw = (tmp_path / "plot" / f"sol-{steps:08d}.vts").read_bytes()
    head, _, rest = raw.partition(b'<AppendedData encoding="raw">')
    head = head.decode()
    assert '<VTKFile type="StructuredGrid"' in head and 'header_type="UInt64"' in head
    assert f'WholeExtent="0 {N - 1} 0 {N - 1} 0 {N - 1}"' in head
    names = re.findall(r'Name="([^"]+)"', head)
    offs = [int(o) for o in re.findall(r'offset="(\d+)"', head)]
    assert names == ["Position", "P", "U", "V", "W"]
    data = rest[rest.index(b"_") + 1:]
    blocks = []
    for o in offs:
        nb = int(np.frombuffer(data[o:o + 8], "<u8")[0])
        blocks.append(np.frombuffer(data[o + 8:o + 8 + nb], "<f8"))
    assert np.max(np.abs(blocks[0].reshape(-1, 3) - want)) < 1e-14
    for f in range(4):
        assert np.array_equal(blocks[1 + f], final[f::4])
