"""f3d_cmass3d_noalloc, the entry point that allocates nothing, against the
cmass3d_ drop-in (Sources/cmass3d.f90) it shares its kernels with: for every
member the four budget terms are the same bits.  Two calls in a row, the
second on a smaller grid with fewer members, so partials the first call left
in the shared scratch must not leak into the second; more members than the
scratch holds are refused."""
import ctypes

import numpy as np
import pytest

from fall3d_amd import lib as F

pytestmark = pytest.mark.gpu

dp = ctypes.POINTER(ctypes.c_double)


def _P(a):
    return a.ctypes.data_as(dp)


def _host_case(rng, nx, ny, nz, nc, nm):
    c = [np.asfortranarray(rng.random((nx + 2, ny + 2, nz + 2, nc)) * 1e-3) for _ in range(nm)]
    dep = [[np.asfortranarray(rng.random((nx, ny, nc + 1))) for _ in range(3)] for _ in range(nm)]
    dz = np.ascontiguousarray(rng.uniform(200.0, 900.0, nz + 1))
    return c, dep, dz


def _expected(L, c, dep, dz, nx, ny, nz, nc, dx, dy):
    out = []
    i = lambda x: ctypes.byref(ctypes.c_int(x))  # noqa: E731
    for cm, (cl, cd, cw) in zip(c, dep):
        o = [ctypes.c_double() for _ in range(4)]
        L.cmass3d_(_P(cm), _P(cl), _P(cd), _P(cw), i(nx), i(ny), i(nz), i(nc),
                   ctypes.byref(ctypes.c_double(dx)), ctypes.byref(ctypes.c_double(dy)), _P(dz),
                   *[ctypes.byref(x) for x in o])
        out.extend(x.value for x in o)
    return np.array(out)


class _Hip:
    """Device buffers through the HIP runtime the library itself links
    (test plumbing: the caller of the entry point owns its memory)."""

    def __init__(self):
        F.lib()  # loads libamdhip64 as a dependency
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def up(self, a):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def down(self, p, n):
        out = np.empty(n)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _noalloc(L, hip, c, dep, dz, nx, ny, nz, nc, dx, dy, stream):
    cd = hip.up(np.concatenate([m.ravel(order="F") for m in c]))
    cl, cdr, cw = (hip.up(np.concatenate([d[q].ravel(order="F") for d in dep])) for q in range(3))
    dzd = hip.up(dz)
    out = hip.up(np.full(4 * len(c), np.nan))
    rc = L.f3d_cmass3d_noalloc(cd, cl, cdr, cw, nx, ny, nz, nc, len(c), dx, dy, dzd, out, stream)
    assert rc == 0, F.last_error()
    assert hip.h.hipStreamSynchronize(stream) == 0
    got = hip.down(out, 4 * len(c))
    hip.free()
    return got


def test_noalloc_cmass3d_matches_the_dropin_on_two_calls():
    L = F.lib()
    hip = _Hip()
    stream = ctypes.c_void_p()
    assert hip.h.hipStreamCreate(ctypes.byref(stream)) == 0
    rng = np.random.default_rng(2024)
    for (nx, ny, nz, nc, nm, dx, dy) in ((37, 29, 11, 3, 5, 5000.0, 4000.0),
                                          (12, 9, 5, 2, 2, 1500.0, 2500.0)):
        c, dep, dz = _host_case(rng, nx, ny, nz, nc, nm)
        want = _expected(L, c, dep, dz, nx, ny, nz, nc, dx, dy)
        got = _noalloc(L, hip, c, dep, dz, nx, ny, nz, nc, dx, dy, stream)
        print(f"[noalloc] {nx}x{ny}x{nz}x{nc}, {nm} members: {got[:4]}", flush=True)
        assert np.isfinite(want).all() and (want > 0).all()
        assert np.array_equal(got, want), (got, want)
    hip.h.hipStreamDestroy(stream)


def test_noalloc_refuses_more_members_than_its_scratch():
    L = F.lib()
    rc = L.f3d_cmass3d_noalloc(1, 1, 1, 1, 4, 4, 4, 1, 65, 1.0, 1.0, 1, 1, None)
    assert rc == -1 and "scratch holds 1 to 64" in F.last_error()
