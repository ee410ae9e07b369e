"""Batched ensemble of independent acoustic problems stepped by one persistent
kernel over a tile work queue (include/acx_ensemble.h)."""
import ctypes as _C

import numpy as np

from . import _check, _ptr, _stream, lib


def _bind():
    L = lib()
    if getattr(L, "_ens_bound", False):
        return L
    vp, i32, dbl = _C.c_void_p, _C.c_int, _C.c_double
    for name, args in {
        "acx_ensemble_create": [i32, _C.POINTER(_C.c_int32), dbl, _C.POINTER(vp)],
        "acx_ensemble_destroy": [vp],
        "acx_ensemble_offsets": [vp, _C.POINTER(_C.c_size_t)],
        "acx_ensemble_copy_in": [vp, vp, vp, vp],
        "acx_ensemble_copy_out": [vp, vp, vp, vp],
        "acx_ensemble_set_config": [vp, i32],
        "acx_ensemble_get_config": [vp],
        "acx_ensemble_step": [vp, dbl, vp],
        "acx_ensemble_graph_steps": [vp, dbl, i32, vp],
        "acx_ensemble_autotune": [vp, dbl, vp, _C.POINTER(i32), _C.POINTER(_C.c_float)],
        "acx_ensemble_num_configs": [],
    }.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = i32
    L._ens_bound = True
    return L


class Ensemble:
    """dims: list of (nx, ny, nz); common cell size h; FP32 compensated state
    packed problem after problem (each a DMDA global vector, 4 floats/cell)."""

    def __init__(self, dims, h):
        L = _bind()
        self.dims = [tuple(int(v) for v in d) for d in dims]
        arr = (_C.c_int32 * (3 * len(dims)))(*[v for d in self.dims for v in d])
        hd = _C.c_void_p()
        _check(L.acx_ensemble_create(len(dims), arr, h, _C.byref(hd)))
        self._h = hd
        self.h = h
        off = (_C.c_size_t * (len(dims) + 1))()
        _check(L.acx_ensemble_offsets(self._h, off))
        self.offsets = list(off)

    def close(self):
        if getattr(self, "_h", None):
            _bind().acx_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ncells(self):
        return self.offsets[-1]

    def load(self, fields64, stream=None):
        """State from a list of FP64 arrays (nz, ny, nx, 4): hi = fp32(U), lo = fp32(U - hi)."""
        import torch
        packed = torch.cat([torch.as_tensor(np.ascontiguousarray(f), dtype=torch.float64).reshape(-1)
                            for f in fields64]).cuda()
        hi = packed.to(torch.float32)
        lo = (packed - hi.to(torch.float64)).to(torch.float32)
        _check(_bind().acx_ensemble_copy_in(self._h, _ptr(hi, torch.float32), _ptr(lo, torch.float32),
                                            _stream(stream)))
        torch.cuda.synchronize()

    def state32(self, stream=None):
        import torch
        n = 4 * self.ncells
        hi = torch.empty(n, dtype=torch.float32, device="cuda")
        lo = torch.empty(n, dtype=torch.float32, device="cuda")
        _check(_bind().acx_ensemble_copy_out(self._h, _ptr(hi, torch.float32), _ptr(lo, torch.float32),
                                             _stream(stream)))
        return hi, lo

    def split(self, flat):
        out = []
        for p, d in enumerate(self.dims):
            a, b = 4 * self.offsets[p], 4 * self.offsets[p + 1]
            out.append(flat[a:b].reshape(d[2], d[1], d[0], 4))
        return out

    def state64(self):
        hi, lo = self.state32()
        return self.split((hi.double() + lo.double()).cpu().numpy())

    def step(self, dt, stream=None):
        _check(_bind().acx_ensemble_step(self._h, dt, _stream(stream)))

    def graph_steps(self, dt, nsteps, stream=None):
        _check(_bind().acx_ensemble_graph_steps(self._h, dt, nsteps, _stream(stream)))

    def set_config(self, c):
        _check(_bind().acx_ensemble_set_config(self._h, c))

    def config(self):
        return _bind().acx_ensemble_get_config(self._h)

    @staticmethod
    def num_configs():
        """Number of tile configurations (set_config / autotune)."""
        return _bind().acx_ensemble_num_configs()

    def autotune(self, dt, stream=None):
        best, ms = _C.c_int(), _C.c_float()
        _check(_bind().acx_ensemble_autotune(self._h, dt, _stream(stream), _C.byref(best), _C.byref(ms)))
        return best.value, ms.value


def random_dims(nprob, seed, sizes=(64, 96, 128, 160)):
    """Edge lengths drawn per problem and axis from `sizes` (seeded)."""
    rng = np.random.default_rng(seed)
    return [tuple(int(v) for v in rng.choice(sizes, 3)) for _ in range(nprob)]
