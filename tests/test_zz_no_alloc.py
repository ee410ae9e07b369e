"""The selective step without allocation (acx_ssp_rks3_step_f32c_sel_noalloc)
against the allocating one (acx_ssp_rks3_step_f32c_sel), through the C ABI.

Both run the same kernels; the no-allocation entry point keeps the tile mask
in a static device array instead of a per-call buffer, so the results must be
equal bit for bit.  Two calls in a row, the second on a smaller grid (the
array keeps the first call's mask, which the second must overwrite), and a
grid with more patch tiles than the array holds is rejected.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flags(torch, grid, seed):
    """A blob of flagged cells (all four component bits) around a seeded
    point, plus a few isolated ones."""
    rng = np.random.default_rng(seed)
    nz, ny, nx, _ = grid.shape()
    f = np.zeros((nz, ny, nx), np.uint8)
    c = rng.integers(0, [nz, ny, nx])
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    d2 = ((z - c[0] + nz // 2) % nz - nz // 2) ** 2 + ((y - c[1] + ny // 2) % ny - ny // 2) ** 2 \
        + ((x - c[2] + nx // 2) % nx - nx // 2) ** 2
    f[(d2 > 36) & (d2 < 64)] = 0xF
    for _ in range(5):
        f[tuple(rng.integers(0, [nz, ny, nx]))] = 0xF
    return torch.from_numpy(f.ravel()).cuda()


def _arrays(torch, acx, grid, seed):
    from acoustic_petsc_amd import learn
    U = torch.from_numpy(learn.steep_pulse_field(grid, seed))
    hi = U.to(torch.float32).cuda()
    lo = (U - hi.cpu().to(torch.float64)).to(torch.float32).cuda()
    return hi, lo


def _step(acx, torch, fn_name, grid, hi, lo, flags, dt):
    L = acx.lib()
    fn = getattr(L, fn_name)
    vp = C.c_void_p
    fn.argtypes = [C.c_void_p, vp, vp, vp, vp, vp, vp, vp, C.c_double, vp]
    fn.restype = C.c_int
    out_hi, out_lo = torch.empty_like(hi), torch.empty_like(lo)
    d0, d1 = torch.empty_like(hi), torch.empty_like(hi)
    g = grid.c_struct()
    rc = fn(C.cast(C.pointer(g), C.c_void_p), vp(hi.data_ptr()), vp(lo.data_ptr()), vp(out_hi.data_ptr()),
            vp(out_lo.data_ptr()), vp(d0.data_ptr()), vp(d1.data_ptr()), vp(flags.data_ptr()), dt,
            acx._stream())
    return rc, out_hi, out_lo


def test_no_alloc_selective_step_equals_the_allocating_one(acx, torch_gpu):
    torch = torch_gpu
    acx.nonfinite_clear()
    for call, (n, seed) in enumerate([((64, 60, 56), 31), ((40, 36, 44), 32)]):
        grid = acx.Grid(n=n, h=1.0 / n[0], system=acx.ACOUSTIC4, gamma_hi=0.85, gamma_lo=0.85)
        print(f"[case] no-alloc call {call}: grid {n} seed {seed}", flush=True)
        hi, lo = _arrays(torch, acx, grid, seed)
        flags = _flags(torch, grid, seed)
        dt = acx.cfl_dt(grid, 0.4)
        rc_a, hi_a, lo_a = _step(acx, torch, "acx_ssp_rks3_step_f32c_sel", grid, hi, lo, flags, dt)
        rc_b, hi_b, lo_b = _step(acx, torch, "acx_ssp_rks3_step_f32c_sel_noalloc", grid, hi, lo, flags, dt)
        torch.cuda.synchronize()
        assert rc_a == 0 and rc_b == 0, acx.lib().acx_last_error()
        assert torch.equal(hi_a, hi_b) and torch.equal(lo_a, lo_b), f"call {call} differs"
        # the step did work: the flagged cells moved away from the linear step
        assert not torch.equal(hi_a, hi)
    acx.nonfinite_check()


def test_no_alloc_selective_step_rejects_grids_beyond_its_scratch(acx, torch_gpu):
    torch = torch_gpu
    L = acx.lib()
    # 30 x 14 x 8-cell patch tiles: 1024^2 x 1024 cells need 35 x 74 x 128 = 331520 > 262144 tiles
    grid = acx.Grid(n=(1024, 1024, 1024), h=1.0 / 1024, system=acx.ACOUSTIC4, gamma_hi=0.85, gamma_lo=0.85)
    g = grid.c_struct()
    assert L.acx_sel_mask_bytes(C.byref(g)) > 5 * 262144 + 4   # a byte and an int32 per tile, and the count
    fn = L.acx_ssp_rks3_step_f32c_sel_noalloc
    fn.argtypes = [C.c_void_p] + [C.c_void_p] * 7 + [C.c_double, C.c_void_p]
    fn.restype = C.c_int
    fake = [C.c_void_p(0x1000 * (k + 1)) for k in range(7)]   # never dereferenced: rejected on the host
    rc = fn(C.cast(C.pointer(g), C.c_void_p), *fake, 1e-4, acx._stream())
    assert rc == 1 and b"static scratch" in L.acx_last_error()
